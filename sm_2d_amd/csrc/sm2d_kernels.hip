// sm2d_kernels.hip — gfx950 kernels of the K1–K6 transport step.
//
// K1  k_mask         ActiveMask per occupied cell (src/cuda/kernels/k1_activemask.cu:6-48)
// K2  k_scan/k_emit  CompactActive: per-block counts, one-workgroup exclusive scan, ordered
//                    scatter with wave ballots -> ascending active / coarse lists
//                    (replaces the atomic compaction of k1k6_pipeline.cu:589-665)
// K3  k_transport    coarse (K2_CoarseTransport, k2_coarsetransport.cu:76-704) and fine
//                    (K3_FineTransport, k3_finetransport.cu:94-760) per-bin transport: one
//                    wavefront per listed cell; lanes evaluate the physics of up to 64 non-zero
//                    bins at once, then the side effects are applied in bin order so slot
//                    claims and FP32 accumulation order equal the reference's serial loop
// K4  k_gather       BucketTransfer (k4_transfer.cu:42-201) as a per-destination gather in
//                    fixed (face, bucket slot) order: one writer per destination cell
// K5  k_audit_*      ConservationAudit (k5_audit.cu:22-167) as fixed-order FP64 reductions
// K6  host           buffer swap; the next iteration resets only the cells that were occupied
#include "sm2d_device.h"
#include "sm2d_mlp.hip"

namespace sm2d {

// Diagnostic build (-DSM2D_STAGE_CLOCKS, python -m sm_2d_amd.build --stage-clocks): every
// transport workgroup sums the constant-rate clock ticks of each stage of its cells in registers
// and its first thread adds the sums to stage_clk[stage] when the workgroup ends (STAGES below;
// [15] counts cells), so the stamps add no memory traffic inside the stages.  The work is the
// same as in the normal build; the normal build compiles the stamps away.
#ifdef SM2D_STAGE_CLOCKS
__device__ unsigned long long stage_clk[32];
// parts of the physics stage ([16 + k]), stamped inside the per-component physics
struct SubClk {
    unsigned long long c[8];
    unsigned long long p;
};
#define SM2D_SUBP , SubClk *sc_
#define SM2D_SUBA(x) , x
#define SM2D_SUB_START()                   \
    do {                                   \
        if (sc_) sc_->p = wall_clock64();  \
    } while (0)
#define SM2D_SUB(k)                                   \
    do {                                              \
        if (sc_) {                                    \
            const unsigned long long t_ = wall_clock64(); \
            sc_->c[k] += t_ - sc_->p;                 \
            sc_->p = t_;                              \
        }                                             \
    } while (0)
#define SM2D_STAMP(k)                                     \
    do {                                                  \
        const unsigned long long t_ = wall_clock64();     \
        clk_acc[k] += t_ - clk_prev;                      \
        clk_prev = t_;                                    \
    } while (0)
#define SM2D_STAMP_START()                                \
    unsigned long long clk_prev = wall_clock64();         \
    unsigned long long clk_acc[12] = {};                  \
    SubClk clk_sub = {};                                  \
    SubClk* sc_ = &clk_sub;                               \
    unsigned long long clk_cells = 0
#define SM2D_STAMP_RESTART() clk_prev = wall_clock64()
#define SM2D_STAMP_CELL() ++clk_cells
#define SM2D_STAMP_FLUSH()                                \
    do {                                                  \
        if (threadIdx.x == 0) {                           \
            for (int k_ = 0; k_ < 12; ++k_)               \
                atomicAdd(&stage_clk[k_], clk_acc[k_]);   \
            for (int k_ = 0; k_ < 8; ++k_)                \
                atomicAdd(&stage_clk[16 + k_], clk_sub.c[k_]); \
            atomicAdd(&stage_clk[15], clk_cells);         \
        }                                                 \
    } while (0)
#else
#define SM2D_STAMP(k) \
    do {              \
    } while (0)
#define SM2D_STAMP_START() \
    do {                   \
    } while (0)
#define SM2D_STAMP_RESTART() \
    do {                     \
    } while (0)
#define SM2D_STAMP_CELL() \
    do {                  \
    } while (0)
#define SM2D_STAMP_FLUSH() \
    do {                   \
    } while (0)
#define SM2D_SUBP
#define SM2D_SUBA(x)
#define SM2D_SUB_START() \
    do {                 \
    } while (0)
#define SM2D_SUB(k) \
    do {            \
    } while (0)
#endif
// stages: 0 tags and destination marks, 1 own psi_out staging, 2 count and queue of non-zero
// bins, 3 physics and per-lane preparation (wave 0's batches), 4 wait for the other waves'
// physics, 5 claims, 6 ordered adds, 7 serial apply, 8 end-of-round barrier, 9 table flush,
// 10 tallies, 11 accumulators.

// Learned straggling width, scattering width and attenuation exponent of a fine-transport
// substep from the surrogate table (linear in ln dR between the two nearest nodes); false when
// the row's material, the step or the table entry is outside the network's training range.
__device__ inline bool surrogate_lookup(const Tabs& T, int ebin, int iz, float dR, float& sE, float& sth,
                                        float& a) {
    if (!T.surr) return false;
    const int m = T.rowmat[iz];
    if (m < 0 || !(dR >= mlp::DR_LO && dR <= mlp::DR_HI)) return false;
    const float lo = logf_(mlp::DR_LO), hi = logf_(mlp::DR_HI);
    const float pj = (logf_(dR) - lo) / (hi - lo) * (float)(T.surr_nodes - 1);
    const int j = min((int)pj, T.surr_nodes - 2);
    const float f = pj - (float)j;
    const float* t0 = T.surr + ((size_t)(m * T.NE + ebin) * T.surr_nodes + j) * 3;
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


// ------------------------------------------------------------------------------------------
// Utility kernels
// ------------------------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Opens an iteration: epoch and iteration counters live on the device so that a chunk of
// iterations can be queued (or replayed from a graph) without host round trips.
__device__ inline void open_iteration(uint32_t* ctl) {
    ctl[C_EP] += 1;
    ctl[C_IT] += 1;
    ctl[C_NDEST] = ctl[C_NDEST1] = 0;
    for (int s = 0; s < 4 * NSH; ++s) ctl[C_BPC + s * SH_STRIDE] = 0;  // bucket pools, destination lists
}

__global__ void k_begin(uint32_t* ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && !ctl[C_DONE]) open_iteration(ctl);
}

// Resets the slot tables of the cells on an occupied-cell list and the buffer's payload pool
// (thread t of nthr).
__device__ inline void clear_cells(Psi p, const uint32_t* __restrict__ occ, uint32_t n, uint32_t t, uint32_t nthr) {
    for (uint32_t i = t; i < n; i += nthr) {
        const uint32_t c = occ[i];
        const uint32_t k = p.ns[c];
        for (uint32_t s = 0; s < k; ++s) p.bid[(size_t)c * KB + s] = EMPTY;
        p.ns[c] = 0;
    }
    for (uint32_t s = t; s < (uint32_t)NSH; s += nthr) p.pcount[s * SH_STRIDE] = 0;
}

// Opens an iteration (as k_begin, from the first thread) and clears the output buffer over its
// occupied-cell list.  Iterations after the first of a run are opened by k_audit_final instead.
__global__ void k_clear(Psi p, const uint32_t* __restrict__ occ, const uint32_t* nocc, uint32_t* ctl) {
    if (ctl[C_DONE]) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) open_iteration(ctl);
    clear_cells(p, occ, *nocc, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// ------------------------------------------------------------------------------------------
// K1: mask + per-cell FP64 totals of psi_in over its occupied-cell list (ascending).  One
// wavefront per occupied cell: the slots' payloads are read coalesced and summed in FP64 with a
// fixed butterfly.  The threshold test W > w_min is defined on the reference's serial FP32 sum
// (slot -> lidx order); that sum differs from the exact sum of n <= 32*LB non-negative terms by
// less than n * 2^-24 relative (< 5e-4), so the FP64 sum decides unless it lies within 1e-3 of
// w_min (or a weight is negative), where lane 0 forms the serial FP32 sum itself.
// Per-entry flag: 1 = active (fine), 2 = coarse (mask 0 with weight), 0 = nothing.
// ------------------------------------------------------------------------------------------
// The FP64 weight and energy of one occupied cell (one wavefront, slots in order, SU slots'
// payloads and bin centres fetched together, positive weights only) and whether a weight is
// negative or NaN (`bad`) or NaN (`nan`).
template <class LY>
__device__ inline void k1_sums(const Tabs& T, const Psi& buf, uint32_t ns, uint32_t mb, uint32_t mr, double& W64,
                               double& E64, bool& bad, bool& nan) {
    const int l = lane_id();
    W64 = 0.0;
    E64 = 0.0;
    bool neg = false, nn = false;
    constexpr int SU = LY::PL >= 8 ? 1 : 8 / LY::PL;
    for (uint32_t s0 = 0; s0 < ns; s0 += SU) {
        float v[SU][LY::PL];
        int bEs[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const uint32_t b = rlu(mb, (int)(s0 + u));  // EMPTY beyond ns
            bEs[u] = b == EMPTY ? -1 : (int)((b >> 12) & 0xFFF);
#pragma unroll
            for (int k = 0; k < LY::PL; ++k) v[u][k] = 0.0f;
            if (b != EMPTY) {
                const float* vp = buf.pool + (size_t)rlu(mr, (int)(s0 + u)) * LY::LB;
#pragma unroll
                for (int k = 0; k < LY::PL; ++k) v[u][k] = vp[l + 64 * k];
            }
        }
        float ec[SU][LY::PL];  // bin centres, fetched without waiting for the weights
#pragma unroll
        for (int u = 0; u < SU; ++u)
#pragma unroll
            for (int k = 0; k < LY::PL; ++k)
                ec[u][k] = ecenter(T.ee, min(max(bEs[u] * LY::EL + LY::le(l + 64 * k), 0), T.NE - 1));
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            if (bEs[u] < 0) continue;
#pragma unroll
            for (int k = 0; k < LY::PL; ++k) {
                const float w = v[u][k];
                if (w > 0.0f) {
                    W64 += (double)w;
                    E64 += (double)w * (double)ec[u][k];
                } else if (w < 0.0f || w != w) {
                    neg = true;
                    nn = nn || (w != w);
                }
            }
        }
    }
    W64 = wave_sum(W64);
    E64 = wave_sum(E64);
    bad = __ballot(neg) != 0;
    nan = __ballot(nn) != 0;
}

// K1 decision for occupied cell c of buffer `buf` in epoch ep (k1_activemask.cu:6-48): any slot
// at or below the on / off energy blocks, the weight test and the hysteresis through the mask of
// epoch ep - 1.  A mask tag is (epoch << 2) | (mask of epoch - 1) << 1 | mask of epoch, so a
// cell whose epoch-ep decision was already made (the previous iteration's closing audit decides
// it ahead, possibly in an earlier call with other thresholds) is decided again from the same
// previous mask with the current thresholds.  Lane 0 writes the tag; returns the list flag (1
// active, 2 coarse, 0 nothing).
template <class LY>
__device__ inline uint8_t k1_decide(const Tp& P, const Psi& buf, uint32_t c, uint32_t ns, uint32_t mb, double W64,
                                    bool bad, uint32_t ep, uint32_t* maskv, uint32_t tag) {
    const int l = lane_id();
    const uint32_t bE = (mb >> 12) & 0xFFF;
    const bool on = __ballot(mb != EMPTY && bE <= (uint32_t)P.b_on) != 0;
    const bool off = __ballot(mb != EMPTY && bE <= (uint32_t)P.b_off) != 0;
    const double wm = (double)P.wmin;
    bool pass;
    if (!bad && W64 > wm * (1.0 + 1e-3)) pass = true;
    else if (!bad && W64 < wm * (1.0 - 1e-3)) pass = false;
    else {
        float W = 0.0f;  // the reference's serial FP32 sum (k1_activemask.cu:24-34)
        if (l == 0)
            for (uint32_t s = 0; s < ns; ++s) {
                const uint32_t b = buf.bid[(size_t)c * KB + s];
                if (b == EMPTY) continue;
                const float* v = buf.pool + (size_t)buf.sref[(size_t)c * KB + s] * LY::LB;
                for (int q = 0; q < LY::LB; ++q) W += v[q];
            }
        pass = rlf(W, 0) > P.wmin;
    }
    uint8_t f = 0;
    if (l == 0 && ns) {
        const uint32_t te = tag >> 2;  // tag: maskv[c], read by the caller before the sums
        const bool prev = (te == ep) ? ((tag & 2u) != 0) : (te == ep - 1u && (tag & 1u));
        const bool m = (on || (prev && off)) && pass;
        maskv[c] = (ep << 2) | (prev ? 2u : 0u) | (m ? 1u : 0u);
        f = m ? 1 : (W64 > 0.0 ? 2 : 0);
    }
    return f;
}

template <class LY>
__global__ __launch_bounds__(64) void k_mask(Tabs T, Tp P, Psi in, const uint32_t* __restrict__ occ,
                                             const uint32_t* nocc, uint32_t* maskv, const uint32_t* ctl,
                                             uint8_t* flag, double* win, double* ein) {
    if (ctl[C_DONE]) return;
    const uint32_t n = *nocc, ep = ctl[C_EP];
    const int l = lane_id();
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t c = occ[i];
        const uint32_t ns = in.ns[c];
        const uint32_t tag = maskv[c];
        const uint32_t mb = (l < (int)ns) ? in.bid[(size_t)c * KB + l] : EMPTY;
        const uint32_t mr = (l < (int)ns) ? in.sref[(size_t)c * KB + l] : 0u;
        double W64, E64;
        bool bad, nan;
        k1_sums<LY>(T, in, ns, mb, mr, W64, E64, bad, nan);
        const uint8_t f = k1_decide<LY>(P, in, c, ns, mb, W64, bad, ep, maskv, tag);
        if (l == 0) {
            flag[i] = f;
            win[i] = W64;
            ein[i] = E64;
        }
    }
}

// Per-tile (256 list entries) counts of the K1 flags, for the ordered compaction.
__global__ __launch_bounds__(256) void k_flag_count(const uint8_t* __restrict__ flag, const uint32_t* nocc,
                                                    uint32_t* tcount, const uint32_t* ctl) {
    if (ctl[C_DONE]) return;
    __shared__ uint32_t s_cnt[2][4];
    const uint32_t n = *nocc;
    const uint32_t ntile = (n + 255) / 256;
    const int w = threadIdx.x >> 6, l = lane_id();
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const uint32_t i = tile * 256 + threadIdx.x;
        const uint8_t f = i < n ? flag[i] : 0;
        const uint64_t ba = __ballot(f == 1), bc = __ballot(f == 2);
        if (l == 0) {
            s_cnt[0][w] = __popcll(ba);
            s_cnt[1][w] = __popcll(bc);
        }
        __syncthreads();
        if (threadIdx.x < 2)
            tcount[tile * 2 + threadIdx.x] =
                s_cnt[threadIdx.x][0] + s_cnt[threadIdx.x][1] + s_cnt[threadIdx.x][2] + s_cnt[threadIdx.x][3];
        __syncthreads();
    }
}

// Exclusive scan of per-tile counts (K counters interleaved) by one workgroup, fixed order.
// With K = 2 (mask compaction) the totals go to ctl[C_NACT], ctl[C_NCRS] and an empty result
// ends the transport (C_DONE, C_DONE_IT); with K = 1 the total goes to *total1.
template <int K>
__global__ __launch_bounds__(1024) void k_scan(uint32_t* tcount, const uint32_t* n_items, uint32_t per_tile,
                                               uint32_t* ctl, uint32_t* total1, uint32_t cap, uint32_t* flags,
                                               uint32_t skip) {
    if (ctl[C_DONE]) {
        // an ended transport has empty lists; a paused one keeps the counts of the iteration that
        // paused it, which size the launches of the chunk queued next
        if (K == 2 && threadIdx.x == 0 && ctl[C_DONE] == 1) {
            ctl[C_NACT] = 0;
            ctl[C_NCRS] = 0;
        }
        return;
    }
    // exclusive scan of 1024 tile counts per pass: wave scans by shuffles, then the 16 wave totals
    __shared__ uint32_t wsum[K][16];
    __shared__ uint32_t carry[K];
    const int ntile = (int)((*n_items + per_tile - 1) / per_tile);
    const int w = (int)(threadIdx.x >> 6), l = lane_id();
    if (threadIdx.x < K) carry[threadIdx.x] = 0;
    __syncthreads();
    for (int base = 0; base < ntile; base += 1024) {
        const int i = base + threadIdx.x;
        uint32_t own[K], inc[K];
        for (int k = 0; k < K; ++k) {
            own[k] = (i < ntile) ? tcount[i * K + k] : 0;
            inc[k] = own[k];
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up((int)inc[k], o);
                if (l >= o) inc[k] += y;
            }
            if (l == 63) wsum[k][w] = inc[k];
        }
        __syncthreads();
        if (w == 0) {
            for (int k = 0; k < K; ++k) {
                uint32_t v = l < 16 ? wsum[k][l] : 0u;  // inclusive scan of the wave totals
                for (int o = 1; o < 16; o <<= 1) {
                    const uint32_t y = (uint32_t)__shfl_up((int)v, o);
                    if (l >= o) v += y;
                }
                if (l < 16) wsum[k][l] = v;
            }
        }
        __syncthreads();
        for (int k = 0; k < K; ++k) {
            const uint32_t before = (w > 0 ? wsum[k][w - 1] : 0u) + inc[k] - own[k];
            if (i < ntile) tcount[i * K + k] = carry[k] + before;
        }
        __syncthreads();
        if (threadIdx.x < K) carry[threadIdx.x] += wsum[threadIdx.x][15];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // A list longer than its buffers stops the transport with a capacity flag, so no later
        // kernel of the queued iterations reads past a buffer.
        bool over = false;
        for (int k = 0; k < K; ++k) over = over || carry[k] > cap;
        if (over) atomicOr(&flags[F_DEST_FULL], 1u);
        if (K == 2) {
            ctl[C_NACT] = over ? 0u : carry[0];
            ctl[C_NCRS] = over ? 0u : carry[1 % K];
            if (over || (carry[0] == 0 && carry[1 % K] == 0)) {
                ctl[C_DONE] = 1;
                ctl[C_DONE_IT] = ctl[C_IT];
            } else if (((skip & 2u) && carry[0]) || ((skip & 1u) && carry[1 % K])) {
                ctl[C_DONE] = 2;  // a list needs a phase this chunk left out: pause before it runs
            }
        } else {
            *total1 = over ? 0u : carry[0];
            if (over) {
                ctl[C_DONE] = 1;
                ctl[C_DONE_IT] = ctl[C_IT];
            }
        }
    }
}

// Ordered scatter of the mask compaction: list entries of a tile in ascending order, ranks
// from wave ballots.  Listed cells are marked occupied for the next iteration's list.
__global__ __launch_bounds__(256) void k_emit(const uint32_t* __restrict__ occ, const uint32_t* nocc,
                                              const uint8_t* flag, const uint32_t* toff, uint32_t* act,
                                              uint32_t* crs, uint32_t* ltag, uint32_t* lpos, uint32_t* bitmap,
                                              const uint32_t* ctl) {
    if (ctl[C_DONE]) return;
    __shared__ uint32_t wcnt[2][4];
    const uint32_t n = *nocc, ep = ctl[C_EP];
    const uint32_t ntile = (n + 255) / 256;
    const int w = threadIdx.x >> 6, l = lane_id();
    const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const uint32_t i = tile * 256 + threadIdx.x;
        const uint8_t f = i < n ? flag[i] : 0;
        const uint64_t ba = __ballot(f == 1), bc = __ballot(f == 2);
        if (l == 0) {
            wcnt[0][w] = __popcll(ba);
            wcnt[1][w] = __popcll(bc);
        }
        __syncthreads();
        if (f) {
            const int k = f - 1;
            const uint32_t c = occ[i];
            uint32_t pos = toff[tile * 2 + k];
            for (int j = 0; j < w; ++j) pos += wcnt[k][j];
            pos += __popcll((k == 0 ? ba : bc) & below);
            if (k == 0) {
                act[pos] = c;
                ltag[c] = 2 * ep + 1;
            } else {
                crs[pos] = c;
                ltag[c] = 2 * ep;
            }
            lpos[c] = pos;
            atomicOr(&bitmap[c >> 5], 1u << (c & 31));
        }
        __syncthreads();
    }
}

// Where the ordered compaction writes (K2 lists, their tags and positions, the occupied bits of
// the listed cells); occ / flag: the occupied list being compacted and its K1 flags.  skip: the
// phases whose kernels the queued chunk leaves out (bit 0 coarse, bit 1 fine); lists that need
// one of them pause the transport (ctl[C_DONE] = 2) before the iteration runs, and the host
// queues the rest again with both phases.
struct ListOut {
    const uint32_t* occ;
    const uint8_t* flag;
    uint32_t *act, *crs, *ltag, *lpos, *bitmap, *flags;
    uint32_t cap;
    uint32_t skip;
};

// One-workgroup ordered compaction of n list entries (k_flag_count + k_scan<2> + k_emit in one
// block of blockDim.x threads, at most 16 waves): tiles in order, ranks from wave ballots,
// running bases per list; lists tagged for epoch ep.  Empty or over-long lists end the
// transport (done_it = the iteration that found them).  wc[2][16], base[2]: shared scratch.
__device__ inline void compact_lists(const ListOut& L, uint32_t n, uint32_t ep, uint32_t done_it, uint32_t* ctl,
                                     uint32_t (*wc)[16], uint32_t* base) {
    const uint32_t nthr = blockDim.x, nw = nthr / 64;
    const int w = threadIdx.x >> 6, l = lane_id();
    const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    if (threadIdx.x < 2) base[threadIdx.x] = 0;
    __syncthreads();
    // Pass 1: list lengths, so that an over-long list writes nothing (as the tiled form).
    uint32_t ca = 0, cc = 0;
    for (uint32_t i = threadIdx.x; i < n; i += nthr) {
        const uint8_t f = L.flag[i];
        ca += f == 1;
        cc += f == 2;
    }
    for (int o = 32; o > 0; o >>= 1) {
        ca += __shfl_xor(ca, o);
        cc += __shfl_xor(cc, o);
    }
    if (l == 0) {
        atomicAdd(&base[0], ca);
        atomicAdd(&base[1], cc);
    }
    __syncthreads();
    const uint32_t na = base[0], nc = base[1];
    __syncthreads();
    if (na > L.cap || nc > L.cap || (na == 0 && nc == 0)) {
        if (threadIdx.x == 0) {
            const bool over = na > L.cap || nc > L.cap;
            if (over) atomicOr(&L.flags[F_DEST_FULL], 1u);
            ctl[C_NACT] = over ? 0u : na;
            ctl[C_NCRS] = over ? 0u : nc;
            ctl[C_DONE] = 1;
            ctl[C_DONE_IT] = done_it;
        }
        return;
    }
    if (threadIdx.x < 2) base[threadIdx.x] = 0;
    __syncthreads();
    // Pass 2: ordered emission.
    for (uint32_t t0 = 0; t0 < n; t0 += nthr) {
        const uint32_t i = t0 + threadIdx.x;
        const uint8_t f = i < n ? L.flag[i] : 0;
        const uint64_t ba = __ballot(f == 1), bc = __ballot(f == 2);
        if (l == 0) {
            wc[0][w] = __popcll(ba);
            wc[1][w] = __popcll(bc);
        }
        __syncthreads();
        if (f) {
            const int k = f - 1;
            uint32_t pos = base[k];
            for (int j = 0; j < w; ++j) pos += wc[k][j];
            pos += __popcll((k == 0 ? ba : bc) & below);
            const uint32_t c = L.occ[i];
            if (k == 0) {
                L.act[pos] = c;
                L.ltag[c] = 2 * ep + 1;
            } else {
                L.crs[pos] = c;
                L.ltag[c] = 2 * ep;
            }
            L.lpos[c] = pos;
            atomicOr(&L.bitmap[c >> 5], 1u << (c & 31));
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            uint32_t s = 0;
            for (uint32_t j = 0; j < nw; ++j) s += wc[threadIdx.x][j];
            base[threadIdx.x] += s;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctl[C_NACT] = na;
        ctl[C_NCRS] = nc;
        if (((L.skip & 2u) && na) || ((L.skip & 1u) && nc)) ctl[C_DONE] = 2;
    }
}

// The compaction as a kernel of its own (1024 threads): the first iteration of a run or of a
// chunk, and lists too long for the closing audit kernel's workgroup.
__global__ __launch_bounds__(1024) void k_compact_small(ListOut L, const uint32_t* nocc, uint32_t* ctl) {
    if (ctl[C_DONE]) {
        if (threadIdx.x == 0 && ctl[C_DONE] == 1) {  // (a pause keeps the counts, as k_scan<2>)
            ctl[C_NACT] = 0;
            ctl[C_NCRS] = 0;
        }
        return;
    }
    __shared__ uint32_t wc[2][16];
    __shared__ uint32_t base[2];
    compact_lists(L, *nocc, ctl[C_EP], ctl[C_IT], ctl, wc, base);
}

// One-workgroup form of k_occ_count + k_scan<1> + k_occ_emit for small grids (nwords <= 16384):
// thread t loads its contiguous run of ceil(nwords / 1024) <= 16 bitmap words once, and one
// workgroup scan of the per-thread popcounts gives its first list position, so the list comes
// out ascending.  An over-long list leaves the bitmap and the list untouched (as the tiled form).
__global__ __launch_bounds__(1024) void k_occ_small(uint32_t* bitmap, uint32_t nwords, uint32_t* occ, uint32_t cap,
                                                    uint32_t* nocc, uint32_t* ctl, uint32_t* flags) {
    if (ctl[C_DONE]) return;
    constexpr int K = 16;
    __shared__ uint32_t sc[16];
    const uint32_t per = (nwords + 1023u) / 1024u;
    const uint32_t w0 = threadIdx.x * per;
    uint32_t wv[K];
    uint32_t cnt = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const uint32_t i = w0 + (uint32_t)j;
        wv[j] = ((uint32_t)j < per && i < nwords) ? bitmap[i] : 0u;
        cnt += __popc(wv[j]);
    }
    const int w = threadIdx.x >> 6, l = lane_id();
    uint32_t x = cnt;  // inclusive scan within the wave
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (l >= o) x += y;
    }
    if (l == 63) sc[w] = x;
    __syncthreads();
    uint32_t pos = x - cnt, total = 0;
    for (int j = 0; j < 16; ++j) {
        const uint32_t v = sc[j];
        if (j < w) pos += v;
        total += v;
    }
    if (total > cap) {
        if (threadIdx.x == 0) {
            atomicOr(&flags[F_DEST_FULL], 1u);
            ctl[C_DONE] = 1;
            ctl[C_DONE_IT] = ctl[C_IT];
            *nocc = 0u;
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
        uint32_t b = wv[j];
        if (b) {
            const uint32_t i = w0 + (uint32_t)j;
            bitmap[i] = 0u;
            while (b) {
                const int k = __ffs(b) - 1;
                b &= b - 1;
                occ[pos] = i * 32 + (uint32_t)k;
                ++pos;
            }
        }
    }
    if (threadIdx.x == 0) *nocc = total;
}

// Occupied-cell list of the buffer just written: per-tile popcounts of the occupancy bitmap
// (256 words = 8192 cells per tile) ...
__global__ __launch_bounds__(256) void k_occ_count(const uint32_t* __restrict__ bitmap, uint32_t nwords,
                                                   uint32_t* tcount, const uint32_t* ctl) {
    if (ctl[C_DONE]) return;
    __shared__ uint32_t wsum[4];
    const uint32_t ntile = (nwords + 255) / 256;
    const int w = (int)(threadIdx.x >> 6), l = lane_id();
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const uint32_t i = tile * 256 + threadIdx.x;
        uint32_t c = i < nwords ? (uint32_t)__popc(bitmap[i]) : 0u;
        for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o);
        if (l == 0) wsum[w] = c;
        __syncthreads();
        if (threadIdx.x == 0) tcount[tile] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
}

// ... and the ordered emission (ascending cells), clearing the bitmap as it goes: positions from
// a wave scan of the words' popcounts and the waves' totals.
__global__ __launch_bounds__(256) void k_occ_emit(uint32_t* bitmap, uint32_t nwords, const uint32_t* toff,
                                                  uint32_t* occ, uint32_t cap, const uint32_t* ctl) {
    if (ctl[C_DONE]) return;
    __shared__ uint32_t wsum[4];
    const uint32_t ntile = (nwords + 255) / 256;
    const int w = (int)(threadIdx.x >> 6), l = lane_id();
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const uint32_t i = tile * 256 + threadIdx.x;
        uint32_t b = i < nwords ? bitmap[i] : 0u;
        const uint32_t c = (uint32_t)__popc(b);
        uint32_t inc = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)inc, o);
            if (l >= o) inc += y;
        }
        if (l == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t before = 0;
        for (int k = 0; k < w; ++k) before += wsum[k];
        uint32_t pos = toff[tile] + before + inc - c;
        if (b) {
            bitmap[i] = 0u;
            while (b) {
                const int k = __ffs(b) - 1;
                b &= b - 1;
                if (pos < cap) occ[pos] = i * 32 + (uint32_t)k;
                ++pos;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// K3 transport
// ------------------------------------------------------------------------------------------
// Finds desired in a 32-slot table held one slot per lane, else claims the next slot, else
// falls back to the closest occupied block (device_psic.cuh:192-260, device_bucket.cuh:61-128).
// Returns the slot (-1 when saturated and empty) and the selected block id.
__device__ inline int table_find_or_claim(uint32_t& mybid, uint32_t& myref, uint32_t& n, uint32_t want,
                                          uint32_t& sel, float* pool, uint32_t* pcount, uint32_t pcap,
                                          uint32_t* gbid, uint32_t* gsref, int LB, uint32_t* flags,
                                          int flag_full) {
    const int l = lane_id();
    sel = want;
    uint64_t hit = __ballot(l < 32 && mybid == want);
    if (hit) return __ffsll((unsigned long long)hit) - 1;
    if (n < (uint32_t)KB) {
        int s = (int)n;
        uint32_t ref = 0;
        if (l == 0) ref = pool_take(pcount, pcap, 1u, blockIdx.x);
        ref = rlu(ref, 0);
        if (ref == POOL_FULL) {
            if (l == 0) atomicOr(&flags[flag_full], 1u);
            return -1;
        }
        for (int i = l; i < LB; i += 64) pool[(size_t)ref * LB + i] = 0.0f;
        if (l == 0) {
            gbid[s] = want;
            gsref[s] = ref;
        }
        if (l == s) {
            mybid = want;
            myref = ref;
        }
        n = n + 1;
        __threadfence_block();
        return s;
    }
    int key = 0x7fffffff;
    if (l < 32 && mybid != EMPTY) {
        int d = abs((int)((mybid >> 12) & 0xFFF) - (int)((want >> 12) & 0xFFF)) * 8 +
                abs((int)(mybid & 0xFFF) - (int)(want & 0xFFF));
        key = (d << 5) | l;
    }
    key = wave_min(key);
    if (key == 0x7fffffff) return -1;
    int s = key & 31;
    sel = rlu(mybid, s);
    return s;
}

template <class LY>
__device__ inline int remap(int l, uint32_t src, uint32_t dst) {  // device_psic.cuh:265-292
    if (src == dst) return l;
    int tg = (int)(src & 0xFFF) * LY::TL + LY::lt(l);
    int eg = (int)((src >> 12) & 0xFFF) * LY::EL + LY::le(l);
    int td = min(max(tg - (int)(dst & 0xFFF) * LY::TL, 0), LY::TL - 1);
    int ed = min(max(eg - (int)((dst >> 12) & 0xFFF) * LY::EL, 0), LY::EL - 1);
    return LY::enc(td, ed, LY::lx(l), LY::lz(l));
}

struct CellAccum {
    double edep, E_cut, E_nuc, E_bnd, drop_w, drop_E, cl_win, cl_wout, cl_Ein, cl_Eout;
    float w_cut, w_nuc, w_bnd;
    uint32_t comps;
};

// Phase-space tables of one listed cell, held by its wavefront: table 0 is the cell's own
// psi_out, tables 1..4 its outflow buckets of faces 0..3.  Lane s < 32 holds slot s (block id,
// payload reference).  New payloads live in a per-wave LDS row pool (reference CACHED | row)
// and are written to the global pools once, when the cell ends; slots beyond the LDS rows and
// the psi_out slots that existed before the cell started use the global pools directly.
constexpr uint32_t CACHED = 0x80000000u;

template <class LY>
struct WTabs {
    uint32_t b0, b1, b2, b3, b4;  // block id of slot (lane)
    uint32_t r0, r1, r2, r3, r4;  // payload reference of slot (lane)
    uint32_t n0, n1, n2, n3, n4;  // slots in use (wave-uniform)
    uint32_t own0;                // own psi_out slots present when the cell started
    uint32_t orig;                // global reference of a preloaded own slot (lane < own0)
    uint32_t used;                // LDS rows in use (wave-uniform)
    uint32_t cap;                 // LDS rows available
    float* rows;                  // this wave's LDS rows [cap][LB]

    // Getters read every field and combine them with masks: a table-dependent branch or address
    // would make the compiler keep the whole struct in scratch memory.
    static __device__ uint32_t pick(int t, uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t a4) {
        const uint32_t m0 = 0u - (uint32_t)(t == 0), m1 = 0u - (uint32_t)(t == 1), m2 = 0u - (uint32_t)(t == 2),
                       m3 = 0u - (uint32_t)(t == 3), m4 = 0u - (uint32_t)(t == 4);
        return (a0 & m0) | (a1 & m1) | (a2 & m2) | (a3 & m3) | (a4 & m4);
    }
    __device__ uint32_t gb(int t) const { return pick(t, b0, b1, b2, b3, b4); }
    __device__ uint32_t gr(int t) const { return pick(t, r0, r1, r2, r3, r4); }
    __device__ uint32_t gn(int t) const { return pick(t, n0, n1, n2, n3, n4); }
    // Setters write every field through a select (no store through a table-dependent address),
    // so the compiler keeps the tables in registers.
    static __device__ uint32_t blend(bool c, uint32_t v, uint32_t old) {
        const uint32_t m = 0u - (uint32_t)c;
        return (v & m) | (old & ~m);
    }
    __device__ void sb(int t, uint32_t v, bool on = true) {
        b0 = blend(on && t == 0, v, b0);
        b1 = blend(on && t == 1, v, b1);
        b2 = blend(on && t == 2, v, b2);
        b3 = blend(on && t == 3, v, b3);
        b4 = blend(on && t == 4, v, b4);
    }
    __device__ void sr(int t, uint32_t v, bool on = true) {
        r0 = blend(on && t == 0, v, r0);
        r1 = blend(on && t == 1, v, r1);
        r2 = blend(on && t == 2, v, r2);
        r3 = blend(on && t == 3, v, r3);
        r4 = blend(on && t == 4, v, r4);
    }
    __device__ void sn(int t, uint32_t v) {
        n0 = blend(t == 0, v, n0);
        n1 = blend(t == 1, v, n1);
        n2 = blend(t == 2, v, n2);
        n3 = blend(t == 3, v, n3);
        n4 = blend(t == 4, v, n4);
    }

    // psi_out slots that already hold weight are staged in LDS rows as well and written back
    // when the cell ends: slot s takes row s while rows remain (used is 0 here).  The payloads
    // of SU slots are fetched together.
    __device__ void stage_own(const Psi& out) {
        const int l = lane_id();
        constexpr int SU = LY::PL >= 8 ? 1 : 8 / LY::PL;
        const uint32_t n = own0 < cap ? own0 : cap;
        for (uint32_t s0 = 0; s0 < n; s0 += SU) {
            float v[SU][LY::PL];
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const uint32_t sl = s0 + (uint32_t)u;
                if (sl < n) {
                    const float* src = out.pool + (size_t)rlu(r0, (int)sl) * LY::LB;
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) v[u][k] = src[l + 64 * k];
                }
            }
#pragma unroll
            for (int u = 0; u < SU; ++u) {
                const uint32_t sl = s0 + (uint32_t)u;
                if (sl < n) {
                    float* row = rows + (size_t)sl * LY::LB;
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) row[l + 64 * k] = v[u][k];
                }
            }
        }
        if ((uint32_t)l < n) {
            orig = r0;
            r0 = CACHED | (uint32_t)l;
        }
        used = n;
    }

    // Same search / claim / closest-block fallback as table_find_or_claim; a claimed slot gets a
    // zeroed LDS row while rows remain, else a zeroed global payload.
    // fallback = false: a full table without `want` returns -2 instead of the closest block.
    __device__ int claim(int t, uint32_t want, uint32_t& sel, float* gpool, uint32_t* gpcount, uint32_t gpcap,
                         uint32_t* flags, int flag_full, bool fallback = true) {
        const int l = lane_id();
        sel = want;
        const uint32_t mb = gb(t);
        const uint64_t hit = __ballot(l < 32 && mb == want);
        if (hit) return __ffsll((unsigned long long)hit) - 1;
        const uint32_t n = gn(t);
        if (n < (uint32_t)KB) {
            const int s = (int)n;
            uint32_t ref;
            if (used < cap) {
                ref = CACHED | used;
                float* row = rows + (size_t)used * LY::LB;
#pragma unroll
                for (int k = 0; k < LY::PL; ++k) row[l + 64 * k] = 0.0f;
                used = used + 1;
            } else {
                ref = 0;
                if (l == 0) ref = pool_take(gpcount, gpcap, 1u, blockIdx.x);
                ref = rlu(ref, 0);
                if (ref == POOL_FULL) {
                    if (l == 0) atomicOr(&flags[flag_full], 1u);
                    return -1;
                }
#pragma unroll
                for (int k = 0; k < LY::PL; ++k) gpool[(size_t)ref * LY::LB + l + 64 * k] = 0.0f;
                __threadfence_block();
            }
            sb(t, want, l == s);
            sr(t, ref, l == s);
            sn(t, n + 1);
            return s;
        }
        if (!fallback) return -2;
        int key = 0x7fffffff;
        if (l < 32 && mb != EMPTY) {
            int d = abs((int)((mb >> 12) & 0xFFF) - (int)((want >> 12) & 0xFFF)) * 8 +
                    abs((int)(mb & 0xFFF) - (int)(want & 0xFFF));
            key = (d << 5) | l;
        }
        key = wave_min(key);
        if (key == 0x7fffffff) return -1;
        const int s = key & 31;
        sel = rlu(mb, s);
        return s;
    }

    // payload[ref][lidx] += w (called by the lanes that add)
    __device__ void add(uint32_t ref, int lidx, float w, float* gpool) const {
        if (ref & CACHED) {
            float* p = rows + (size_t)(ref & ~CACHED) * LY::LB + lidx;
            *p += w;
        } else {
            gpool[(size_t)ref * LY::LB + lidx] += w;
        }
    }
};

// Angular split and energy bin of an emission (device_emit_component_to_bucket_4d,
// device_bucket.cuh:431-533): depends only on the emitted (theta, E), so it is evaluated in the
// lane-parallel phase.
struct Bins {
    int bc, tb1, eb;
    float tw0, tw1, Ecen;
};

__device__ inline Bins emission_bins(const Tabs& T, float theta, float E, int eguess = -1) {
    const int NT = T.NT, NE = T.NE;
    const float tmin = T.te[0], tmax = T.te[NT];
    const float dth = (tmax - tmin) / NT;
    Bins b;
    float tc = fmaxf(tmin, fminf(theta, tmax));
    float c0 = tmin + 0.5f * dth;
    int bc = (int)floorf((tc - c0) / dth + 0.5f);
    bc = max(0, min(bc, NT - 1));
    float delta = tc - (c0 + (float)bc * dth);
    int tb1 = bc;
    float tw0 = 1.0f, tw1 = 0.0f;
    if (delta > 0.0f && bc < NT - 1) tb1 = bc + 1;
    else if (delta < 0.0f && bc > 0) tb1 = bc - 1;
    if (tb1 != bc) {
        float fr = fminf(1.0f, fabsf(delta) / dth);
        tw1 = fr * fr;
        tw0 = 1.0f - tw1;
    }
    float Ec = fmaxf(T.ee[0], fminf(E, T.ee[NE]));
    int eb = find_bin(T.ee, NE, Ec, eguess);
    eb = max(0, min(eb, NE - 1));
    b.bc = bc;
    b.tb1 = tb1;
    b.tw0 = tw0;
    b.tw1 = tw1;
    b.eb = eb;
    b.Ecen = ecenter(T.ee, eb);
    return b;
}

// Wave-uniform broadcast of one component's results (lane j -> all lanes).
struct Comp {
    int kind;     // 0 prune, 1 cutoff before the step, 2 transported
    uint32_t fmask;  // face 0: lateral targets q (bit q) that are other listed cells
    int surr;     // 1: the learned-physics table gave the substep's widths
    float sdev;   // monitored cells: deviation of the learned widths from the plain ones
    float w, E;
    float edep, w_rem, E_rem, w_new, E_new;
    int face;     // -1 stays in the cell
    float th_new, x_new, z_new;
    float sw[11]; // lateral weights (+z exit) or sub-bin fractions (remain)
    float wcf, wl, wr;
    int lidx;
    Bins bn;      // emission bins of (th_new, E_new)
    // face >= 0: g0 = E_emit, g1 = lateral drop with a zero centre drop, g2 = lateral boundary
    //            weight (face 0); i0, i1 = x_sub, z_sub of the emission (faces 1..3)
    // remain:    i0 = theta bin, i1 = E bin, i2 = z_sub, g0 = E bin centre
    int i0, i1, i2;
    float g0, g1, g2;
};

template <class LY>
__device__ inline float emit(WTabs<LY>& W, CellAccum& a, const Tabs& T, Bkt& bk, uint32_t* flags, int face,
                             const Bins& b, float weight, int xs, int zs) {
    if (weight <= 0.0f) return 0.0f;
    const int el = b.eb % LY::EL;
    const uint32_t bE = (uint32_t)(b.eb / LY::EL);
    float dropped = 0.0f;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
        const float tw = ti == 0 ? b.tw0 : b.tw1;
        if (tw <= 0.0f) continue;
        const float we = weight * tw;
        const int t = ti == 0 ? b.bc : b.tb1;
        const uint32_t bid = ((uint32_t)(t / LY::TL) & 0xFFF) | ((bE & 0xFFF) << 12);
        const int lidx = LY::enc(t % LY::TL, el, xs, zs);
        uint32_t sel;
        const int s = W.claim(1 + face, bid, sel, bk.pool, bk.pcount, bk.pcap, flags, F_BKT_FULL);
        if (s < 0) {
            dropped += we;
            continue;
        }
        const int l2 = (sel == bid) ? lidx : remap<LY>(lidx, bid, sel);
        const uint32_t ref = rlu(W.gr(1 + face), s);
        if (lane_id() == 0) W.add(ref, l2, we, bk.pool);
        a.cl_wout += we;
        float Ew = b.Ecen;
        if (sel != bid) {
            const int ebw = min(max((int)((sel >> 12) & 0xFFF) * LY::EL + LY::le(l2), 0), T.NE - 1);
            Ew = ecenter(T.ee, ebw);
        }
        a.cl_Eout += (double)we * (double)Ew;
    }
    return dropped;
}

// End of a cell: one pool reservation per pool for the cell's LDS-cached slots, payload copies
// LDS -> global (coalesced) and the table writes.  A pool overflow raises its flag and leaves
// the affected tables without the new slots, so no later kernel reads past a pool.
// Copies whole LDS payload rows to pool rows, 16 bytes per lane, so that one wavefront
// instruction moves 64 / (LB / 4) rows (two 512 B rows, or one 1 KiB row): the table slots in
// mask m, where lane s holds slot s's LDS row in xr (CACHED | row) and its pool row in yr.
// Fewer, wider stores keep a cell's flush well inside the wave's 63 outstanding memory
// operations.
template <class LY>
__device__ inline void rows_to_pool(uint64_t m, uint32_t xr, uint32_t yr, const float* rows, float* pool) {
    constexpr int LPR = LY::LB / 4;  // lanes per row
    constexpr int RPI = 64 / LPR;    // rows per instruction
    static_assert(LY::LB % 4 == 0 && 64 % LPR == 0, "rows must split into 16-byte lane parts");
    const int l = lane_id(), g = l / LPR, e = l - g * LPR;
    while (m) {
        int s = -1;
#pragma unroll
        for (int k = 0; k < RPI; ++k) {
            if (m) {
                const int sk = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                if (g == k) s = sk;
            }
        }
        const uint32_t x = (uint32_t)__shfl((int)xr, s < 0 ? 0 : s);
        const uint32_t y = (uint32_t)__shfl((int)yr, s < 0 ? 0 : s);
        if (s >= 0) {
            const float4 v = *reinterpret_cast<const float4*>(rows + (size_t)(x & ~CACHED) * LY::LB + 4 * e);
            *reinterpret_cast<float4*>(pool + (size_t)y * LY::LB + 4 * e) = v;
        }
    }
}

// Size of the claimed-key cache of fast_claims (entries) and its hash.
constexpr int NKC = 32;
__device__ inline uint32_t kc_slot(uint32_t K) { return (K * 2654435761u) >> 27; }

// Home pool rows of a transport cell's slots (LDS, table t slot s at home[t * KB + s]): the pool
// row a slot's payload was moved to by evict_rows, or read from by a reload, or (psi_out slots
// that held weight when the cell started) its original payload; NOHOME otherwise.
constexpr uint32_t NOHOME = 0xFFFFFFFFu;

// End of a transport cell: LDS rows go to their home rows, rows without one to rows reserved now
// (one sharded reservation per pool); then the psi_out table and the four bucket tables.
template <class LY>
__device__ inline void flush_tables(WTabs<LY>& W, Psi& out, Bkt& bk, int cell, uint32_t li, const uint32_t* home,
                                    uint32_t* flags) {
    const int l = lane_id();
    const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    uint64_t mc[5], mn[5];
    uint32_t hh[5];
    uint32_t co = 0, cb = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        hh[t] = l < 32 ? home[t * KB + l] : NOHOME;
        const bool c = l < 32 && (uint32_t)l < W.gn(t) && (W.gr(t) & CACHED);
        mc[t] = __ballot(c);
        mn[t] = __ballot(c && hh[t] == NOHOME);
        if (t == 0) co = (uint32_t)__popcll(mn[0]);
        else cb += (uint32_t)__popcll(mn[t]);
    }
    // both pool reservations in one pass: lane 0 psi_out's, lane 1 the buckets'
    uint32_t take = 0;
    if ((l == 0 && co) || (l == 1 && cb))
        take = pool_take(l == 0 ? out.pcount : bk.pcount, l == 0 ? out.pcap : bk.pcap, l == 0 ? co : cb,
                         blockIdx.x);
    const uint32_t base = rlu(take, 0), bbase = rlu(take, 1);
    bool ok = !co || base != POOL_FULL;
    if (!ok && l == 0) atomicOr(&flags[F_PSI_FULL], 1u);
    if (ok) {
        uint32_t dst = hh[0];
        if ((mn[0] >> l) & 1ull) dst = base + (uint32_t)__popcll(mn[0] & below);
        rows_to_pool<LY>(mc[0], W.r0, dst, W.rows, out.pool);
        const uint32_t ref = ((mc[0] >> l) & 1ull) ? dst : W.r0;
        // slots that held weight keep their payload reference
        if (l < 32 && (uint32_t)l >= W.own0 && (uint32_t)l < W.n0) {
            out.bid[(size_t)cell * KB + l] = W.b0;
            out.sref[(size_t)cell * KB + l] = ref;
        }
        if (l == 0) out.ns[cell] = W.n0;
    }
    // buckets of faces 0..3
    ok = !cb || bbase != POOL_FULL;
    if (!ok && l == 0) atomicOr(&flags[F_BKT_FULL], 1u);
    uint32_t k0 = 0;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const uint32_t nf = ok ? W.gn(1 + f) : 0u;
        if (ok) {
            uint32_t dst = hh[1 + f];
            if ((mn[1 + f] >> l) & 1ull) dst = bbase + k0 + (uint32_t)__popcll(mn[1 + f] & below);
            rows_to_pool<LY>(mc[1 + f], W.gr(1 + f), dst, W.rows, bk.pool);
            const uint32_t ref = ((mc[1 + f] >> l) & 1ull) ? dst : W.gr(1 + f);
            const uint32_t base_t = (li * 4 + f) * KB;
            if (l < 32 && (uint32_t)l < nf) {
                bk.bid[base_t + l] = W.gb(1 + f);
                bk.sref[base_t + l] = ref;
            }
            k0 += (uint32_t)__popcll(mn[1 + f]);
        }
        if (l == 0) bk.n[li * 4 + f] = nf;
    }
}

// Mid-cell eviction, when a batch's new keys would need more LDS rows than remain: every LDS row
// goes to its slot's home row (reserved now where there is none), the slots then refer to those
// pool rows, and the LDS rows are free again; fast_claims copies a slot back into an LDS row when
// a later batch adds to it.  The payload values move unchanged, so the adds keep their order and
// their results.  Returns false (with the pool's full flag) when a reservation fails.
template <class LY>
__device__ inline bool evict_rows(WTabs<LY>& W, Psi& out, Bkt& bk, uint32_t* home, uint32_t* kc, uint32_t* flags) {
    const int l = lane_id();
    const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    uint64_t mc[5], mn[5];
    uint32_t hh[5];
    uint32_t co = 0, cb = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        hh[t] = l < 32 ? home[t * KB + l] : NOHOME;
        const bool c = l < 32 && (uint32_t)l < W.gn(t) && (W.gr(t) & CACHED);
        mc[t] = __ballot(c);
        mn[t] = __ballot(c && hh[t] == NOHOME);
        if (t == 0) co = (uint32_t)__popcll(mn[0]);
        else cb += (uint32_t)__popcll(mn[t]);
    }
    uint32_t take = 0;
    if ((l == 0 && co) || (l == 1 && cb))
        take = pool_take(l == 0 ? out.pcount : bk.pcount, l == 0 ? out.pcap : bk.pcap, l == 0 ? co : cb,
                         blockIdx.x);
    const uint32_t base = rlu(take, 0), bbase = rlu(take, 1);
    if ((co && base == POOL_FULL) || (cb && bbase == POOL_FULL)) {
        if (l == 0) atomicOr(&flags[co && base == POOL_FULL ? F_PSI_FULL : F_BKT_FULL], 1u);
        return false;
    }
    uint32_t k0 = 0;
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        uint32_t dst = hh[t];
        if ((mn[t] >> l) & 1ull) dst = (t == 0 ? base : bbase + k0) + (uint32_t)__popcll(mn[t] & below);
        rows_to_pool<LY>(mc[t], W.gr(t), dst, W.rows, t == 0 ? out.pool : bk.pool);
        const bool mine = (mc[t] >> l) & 1ull;
        if (mine) home[t * KB + l] = dst;
        W.sr(t, dst, mine);
        if (t > 0) k0 += (uint32_t)__popcll(mn[t]);
    }
    W.used = 0;
    if (l < NKC) kc[l] = EMPTY;  // the key cache maps keys to LDS rows
    __threadfence_block();        // the pool rows are complete before a reload reads one back
    return true;
}

// ------------------------------------------------------------------------------------------
// Lane-parallel form of the ordered apply.  Each lane turns its component into
//  - claim keys (table << 24 | block id) in the serial claim order,
//  - payload records (key, local bin, weight) whose destinations are distinct within the
//    component, and
//  - per-accumulator contributions (its sums from zero, in the serial order).
// The wave then claims the distinct keys in order of first appearance (= the serial claim
// order) and applies the records in rounds in which every destination takes its next record in
// component order, so every FP32 payload sum is formed in the serial order.  The accumulators
// (tallies, drops, closure sums) take one sum per batch of 64 components: a fixed butterfly
// over the batch's lanes (FP64, FP32 for the three FP32 accumulators), added in batch order.
// A batch whose claims need the closest-block fallback, a global payload or an exhausted pool
// is applied by the serial code instead, with the same batch sums.
// ------------------------------------------------------------------------------------------
enum {
    K_EDEP, K_ECUT, K_ENUC, K_EBND, K_WCUT, K_WNUC, K_WBND,
    K_DROPW, K_DROPE, K_CLWIN, K_CLWOUT, K_CLEIN, K_CLEOUT, NKACC
};
constexpr int MAXKEY = 5;   // own block + two two-bin side emissions

__device__ inline bool kacc_is_float(int a) { return a == K_WCUT || a == K_WNUC || a == K_WBND; }

// LDS scratch of one transport wave for one chunk of 64 components.
template <class LY>
struct FastLds {
    static constexpr int MR = LY::XS + 4;  // records per lane: x_sub spread + two two-bin side emissions
    uint16_t rec[64 * MR];   // (key index << 12 | lidx), then the LDS payload address (< 6144)
    float recw[64 * MR];
    uint32_t key[64 * MAXKEY];  // claim keys; the claims reuse it for the new keys in order
    uint8_t nkey[64], nrec[64];
};

struct LaneOut {
    int nkey, nrec;
    double clwout, cleout;
    CellAccum c;  // the component's accumulator contributions
};

// Lane i receives lane (i ^ M)'s 32-bit value: DPP quad permutations for M = 1, 2, LDS-free
// swizzles for M = 4, 8, 16, a permute for M = 32.  Every lane of the wave must be active.
template <int M>
__device__ inline int lane_xor(int v) {
    if constexpr (M == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // [1,0,3,2]
    else if constexpr (M == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // [2,3,0,1]
    else if constexpr (M == 32) return __shfl_xor(v, 32);
    else return __builtin_amdgcn_ds_swizzle(v, (M << 10) | 0x1F);  // and 0x1F, xor M (32-lane groups)
}
template <int M>
__device__ inline double lane_xor(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)lane_xor<M>((int)(uint32_t)u), hi = (uint32_t)lane_xor<M>((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <int M>
__device__ inline float lane_xor(float v) {
    return __int_as_float(lane_xor<M>(__float_as_int(v)));
}

// Sum over a batch's 64 lanes of one accumulator's contributions, in a fixed butterfly order
// (v_i + v_(i^o) for o = 32, 16, ..., 1; every lane ends with the same sum).  Full wave only.
template <class F>
__device__ inline F bfly_sum(F v) {
    v += lane_xor<32>(v);
    v += lane_xor<16>(v);
    v += lane_xor<8>(v);
    v += lane_xor<4>(v);
    v += lane_xor<2>(v);
    v += lane_xor<1>(v);
    return v;
}

// The batch sums of all accumulators; lane k < NKACC returns accumulator k's sum (the FP32
// sums widened, which is exact).
__device__ inline double batch_sum_lane(const CellAccum& c) {
    // thirteen independent butterflies, interleaved by the compiler
    const double edep = bfly_sum(c.edep), E_cut = bfly_sum(c.E_cut), E_nuc = bfly_sum(c.E_nuc),
                 E_bnd = bfly_sum(c.E_bnd), drop_w = bfly_sum(c.drop_w), drop_E = bfly_sum(c.drop_E),
                 cl_win = bfly_sum(c.cl_win), cl_wout = bfly_sum(c.cl_wout), cl_Ein = bfly_sum(c.cl_Ein),
                 cl_Eout = bfly_sum(c.cl_Eout);
    const float w_cut = bfly_sum(c.w_cut), w_nuc = bfly_sum(c.w_nuc), w_bnd = bfly_sum(c.w_bnd);
    const int l = lane_id();
    return l == K_EDEP ? edep : l == K_ECUT ? E_cut : l == K_ENUC ? E_nuc : l == K_EBND ? E_bnd
         : l == K_WCUT ? (double)w_cut : l == K_WNUC ? (double)w_nuc : l == K_WBND ? (double)w_bnd
         : l == K_DROPW ? drop_w : l == K_DROPE ? drop_E : l == K_CLWIN ? cl_win : l == K_CLWOUT ? cl_wout
         : l == K_CLEIN ? cl_Ein : l == K_CLEOUT ? cl_Eout : 0.0;
}

// Lanes 0..NKACC-1 hold the cell's accumulators (FP64 in accd, FP32 in accf) and add their
// batch sum s.
__device__ inline void add_batch_sum(double s, double& accd, float& accf) {
    const int l = lane_id();
    if (l >= NKACC) return;
    if (kacc_is_float(l)) accf = accf + (float)s;
    else accd += s;
}

// Weight and energy of a component's lateral emissions into other listed cells' +z buckets,
// booked to the source cell when emitted (they are applied after the phase, apply_foreign): per
// target q in fmask and per theta part, the weight (w_new sw[q]) tw and its energy at the
// emission bin's centre.
__device__ inline void foreign_booking(const Comp& r, double& clw, double& cle) {
#pragma unroll
    for (int q = 0; q < 11; ++q) {
        if (!((r.fmask >> q) & 1u)) continue;
        const float wl = r.w_new * r.sw[q];
        if (r.bn.tw0 > 0.0f) {
            const float we = wl * r.bn.tw0;
            clw += (double)we;
            cle += (double)we * (double)r.bn.Ecen;
        }
        if (r.bn.tw1 > 0.0f) {
            const float we = wl * r.bn.tw1;
            clw += (double)we;
            cle += (double)we * (double)r.bn.Ecen;
        }
    }
}

template <class LY>
__device__ inline void rec_emit(FastLds<LY>& F, LaneOut& o, int t, const Bins& b, float weight, int xs, int zs) {
    if (weight <= 0.0f) return;
    const int l = lane_id();
    const int el = b.eb % LY::EL;
    const uint32_t bE = (uint32_t)(b.eb / LY::EL);
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
        const float tw = ti == 0 ? b.tw0 : b.tw1;
        if (tw <= 0.0f) continue;
        const float we = weight * tw;
        const int tb = ti == 0 ? b.bc : b.tb1;
        const uint32_t bid = ((uint32_t)(tb / LY::TL) & 0xFFF) | ((bE & 0xFFF) << 12);
        const int lidx = LY::enc(tb % LY::TL, el, xs, zs);
        const int k = o.nkey++;
        F.key[l * MAXKEY + k] = ((uint32_t)t << 24) | bid;
        const int r = o.nrec++;
        F.rec[l * FastLds<LY>::MR + r] = (uint16_t)(((uint32_t)k << 12) | (uint32_t)lidx);
        F.recw[l * FastLds<LY>::MR + r] = we;
        o.clwout += we;
        o.cleout += (double)we * (double)b.Ecen;
    }
}

// Adds one contribution v to accumulator K of the lane's component (FP64 accumulators widen it).
template <int K>
__device__ inline void cadd(CellAccum& c, float v) {
    if (K == K_EDEP) c.edep += (double)v;
    else if (K == K_ECUT) c.E_cut += (double)v;
    else if (K == K_ENUC) c.E_nuc += (double)v;
    else if (K == K_EBND) c.E_bnd += (double)v;
    else if (K == K_WCUT) c.w_cut += v;
    else if (K == K_WNUC) c.w_nuc += v;
    else if (K == K_WBND) c.w_bnd += v;
    else if (K == K_DROPW) c.drop_w += (double)v;
    else if (K == K_DROPE) c.drop_E += (double)v;
}

// The component of this lane (see the serial apply for the meaning of each quantity).
template <class LY, bool FINE>
__device__ inline LaneOut fast_prepare(FastLds<LY>& F, const Comp& r, bool live, int ix, int iz, int Nx, int Nz) {
    const int l = lane_id();
    LaneOut o;
    o.nkey = o.nrec = 0;
    o.clwout = o.cleout = 0.0;
    o.c = CellAccum{};
    if (!live) return o;
    const float w = r.w, E = r.E;
    o.c.cl_win = (double)w;
    o.c.cl_Ein = (double)w * (double)E;
    if (r.kind == 0) {
        cadd<K_DROPW>(o.c, w);
        cadd<K_DROPE>(o.c, E * w);
        o.clwout = (double)w;
        o.cleout = (double)(E * w);
    } else if (r.kind == 1) {
        cadd<K_WCUT>(o.c, w);
        cadd<K_ECUT>(o.c, E * w);
        o.clwout = (double)w;
        o.cleout = (double)(E * w);
    } else {
        const float w_new = r.w_new, E_new = r.E_new;
        cadd<K_EDEP>(o.c, r.edep);
        cadd<K_WNUC>(o.c, r.w_rem);
        cadd<K_ENUC>(o.c, r.E_rem);
        o.clwout += (double)r.w_rem;
        o.cleout += (double)r.edep + (double)r.E_rem;
        auto bnd = [&](float wb) {
            cadd<K_WBND>(o.c, wb);
            cadd<K_EBND>(o.c, E_new * wb);
            o.clwout += (double)wb;
            o.cleout += (double)(E_new * wb);
        };
        if (E_new <= 0.1f) {
            cadd<K_WCUT>(o.c, w_new);
            cadd<K_ECUT>(o.c, E_new * w_new);
            o.clwout += (double)w_new;
            o.cleout += (double)(E_new * w_new);
        } else if (r.face >= 0) {
            const int face = r.face;
            const bool nb = (face == 0) ? (iz + 1 < Nz) : (face == 1) ? (iz > 0) : (face == 2) ? (ix + 1 < Nx) : (ix > 0);
            if (!nb) {
                bnd(w_new);
            } else if (face == 0) {
                const float E_emit = r.g0, drop = r.g1, lat_bnd = r.g2;
                const float wf5 = r.sw[5];
                if (wf5 > 0.0f) rec_emit<LY>(F, o, 1, r.bn, w_new * wf5, LY::XS / 2, 0);
                foreign_booking(r, o.clwout, o.cleout);
                if (lat_bnd > 0.0f) bnd(lat_bnd);
                if (drop > 0.0f) {
                    cadd<K_DROPW>(o.c, drop);
                    cadd<K_DROPE>(o.c, E_new * drop);
                    o.clwout += (double)drop;
                    o.cleout += (double)(E_new * drop);
                }
                if (FINE) {
                    const float tw = fmaxf(0.0f, w_new - lat_bnd - drop);
                    if (tw > 0.0f) {
                        cadd<K_EDEP>(o.c, (E_new - E_emit) * tw);
                        o.cleout += (double)((E_new - E_emit) * tw);
                    }
                }
            } else {
                const float E_emit = r.g0;
                rec_emit<LY>(F, o, 1 + face, r.bn, w_new, r.i0, r.i1);
                if (FINE) {
                    const float tw = fmaxf(0.0f, w_new - 0.0f);
                    if (tw > 0.0f) {
                        cadd<K_EDEP>(o.c, (E_new - E_emit) * tw);
                        o.cleout += (double)((E_new - E_emit) * tw);
                    }
                }
            }
        } else {
            const int tbn = r.i0, ebn = r.i1, zst = r.i2;
            const float E_emit = r.g0;
            const uint32_t bidn = ((uint32_t)(ebn / LY::EL) << 12) | ((uint32_t)(tbn / LY::TL) & 0xFFF);
            F.key[l * MAXKEY + 0] = bidn;  // table 0
            o.nkey = 1;
#pragma unroll
            for (int q = 0; q < LY::XS; ++q) {
                const float wf = r.sw[q];
                if (wf > 0.0f) {
                    const float wsp = w_new * wf;
                    o.clwout += (double)wsp;
                    o.cleout += (double)wsp * (double)E_emit;
                    const int rr = o.nrec++;
                    F.rec[l * FastLds<LY>::MR + rr] = (uint16_t)LY::enc(tbn % LY::TL, ebn % LY::EL, q, zst);
                    F.recw[l * FastLds<LY>::MR + rr] = wsp;
                }
            }
            const float wcf = r.wcf, wl = r.wl, wr = r.wr;
            if (wcf < 1.0f - 1e-6f) {
                if (wl > 0.0f) {
                    const float wsl = w_new * wl;
                    if (ix > 0) rec_emit<LY>(F, o, 4, r.bn, wsl, LY::XS - 1, zst);
                    else bnd(wsl);
                }
                if (wr > 0.0f) {
                    const float wsr = w_new * wr;
                    if (ix < Nx - 1) rec_emit<LY>(F, o, 3, r.bn, wsr, 0, zst);
                    else bnd(wsr);
                }
            }
            if (FINE) {
                float rf = wcf;
                if (ix > 0) rf += wl;
                if (ix < Nx - 1) rf += wr;
                rf = fminf(1.0f, fmaxf(0.0f, rf));
                const float rw = w_new * rf;
                if (rw > 0.0f) {
                    cadd<K_EDEP>(o.c, (E_new - E_emit) * rw);
                    o.cleout += (double)((E_new - E_emit) * rw);
                }
            }
        }
    }
    o.c.cl_wout = o.clwout;
    o.c.cl_Eout = o.cleout;
    return o;
}

// Claims of the batch's distinct keys in order of first appearance, with the serial claims'
// outcome: a key found in its table takes that slot, any other key takes its table's next free
// slot and the next LDS row.  A found slot whose payload is a pool row (moved out by
// evict_rows, or claimed by the serial path) is copied back into the next LDS row, the pool row
// becoming its home.  The loop over distinct keys only decides; the tables and rows change once
// every key is decided and the capacities are checked, so a return other than FC_OK leaves them
// untouched: FC_ROWS (the LDS rows would run out: evict and try again) or FC_SERIAL (a table
// would overflow into the closest-block fallback: the serial path).  Resolves record addresses.
// Direct-mapped cache of the current cell's claimed keys (key -> LDS row) in 2 x NKC words of
// LDS: keys found in it skip the serial loop over distinct keys.  Only slots whose payload is
// an LDS row enter it, and evict_rows clears it, so a hit gives the loop's answer.
enum { FC_SERIAL = 0, FC_OK = 1, FC_ROWS = 2 };
constexpr uint32_t RELOAD = 0x80000000u;  // pending-key flag: a found slot copied back into LDS

// One attempt with the lane's keys k0..k4 in registers: the attempt reuses F.key for its list
// of pending keys, so the lane's keys are read from F.key once, before the first attempt.
template <class LY>
__device__ inline int fast_claims_try(FastLds<LY>& F, WTabs<LY>& W, uint32_t* kc, uint32_t* home,
                                      const float* psi_pool, const float* bk_pool, int nkey, int nrec, uint32_t k0,
                                      uint32_t k1, uint32_t k2, uint32_t k3, uint32_t k4) {
    const int l = lane_id();
    uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0;
    uint32_t pend = (nkey > 0) ? ((1u << nkey) - 1u) : 0u;
    {
        const uint32_t h0 = kc_slot(k0), h1 = kc_slot(k1), h2 = kc_slot(k2), h3 = kc_slot(k3), h4 = kc_slot(k4);
        const uint32_t c0 = kc[h0], c1 = kc[h1], c2 = kc[h2], c3 = kc[h3], c4 = kc[h4];
        const uint32_t q0 = kc[NKC + h0], q1 = kc[NKC + h1], q2 = kc[NKC + h2], q3 = kc[NKC + h3],
                       q4 = kc[NKC + h4];
        if ((pend & 1u) && c0 == k0) { r0 = q0; pend &= ~1u; }
        if ((pend & 2u) && c1 == k1) { r1 = q1; pend &= ~2u; }
        if ((pend & 4u) && c2 == k2) { r2 = q2; pend &= ~4u; }
        if ((pend & 8u) && c3 == k3) { r3 = q3; pend &= ~8u; }
        if ((pend & 16u) && c4 == k4) { r4 = q4; pend &= ~16u; }
    }
    const uint32_t used0 = W.used;
    uint32_t nnew = 0;  // new slots so far, in order of first appearance (wave-uniform)
    uint32_t cnt = 0;   // new slots per table, 6 bits per table (wave-uniform)
    while (true) {
        const uint64_t hp = __ballot(pend != 0u);
        if (!hp) break;
        const int j = __ffsll((unsigned long long)hp) - 1;
        const int kk = __ffs(rlu(pend, j)) - 1;
        const uint32_t mine = kk == 0 ? k0 : kk == 1 ? k1 : kk == 2 ? k2 : kk == 3 ? k3 : k4;
        const uint32_t K = rlu(mine, j);
        uint32_t hit = 0;
        hit |= (k0 == K) ? 1u : 0u;
        hit |= (k1 == K) ? 2u : 0u;
        hit |= (k2 == K) ? 4u : 0u;
        hit |= (k3 == K) ? 8u : 0u;
        hit |= (k4 == K) ? 16u : 0u;
        hit &= pend;
        pend &= ~hit;
        const int t = (int)(K >> 24);
        const uint64_t found = __ballot(l < 32 && W.gb(t) == (K & 0xFFFFFFu));
        uint32_t row;
        if (found) {
            const uint32_t ref = rlu(W.gr(t), __ffsll((unsigned long long)found) - 1);
            if (ref & CACHED) {
                row = ref & ~CACHED;
                if (l == 0) {
                    const uint32_t h = kc_slot(K);
                    kc[h] = K;
                    kc[NKC + h] = row;
                }
            } else {  // copied back into the next LDS row at the commit
                if (l == 0 && nnew < 64 * MAXKEY) F.key[nnew] = K | RELOAD;
                row = used0 + nnew;
                ++nnew;
            }
        } else {
            if (l == 0 && nnew < 64 * MAXKEY) F.key[nnew] = K;
            row = used0 + nnew;
            ++nnew;
            cnt += 1u << (6 * t);
        }
        if (hit & 1u) r0 = row;
        if (hit & 2u) r1 = row;
        if (hit & 4u) r2 = row;
        if (hit & 8u) r3 = row;
        if (hit & 16u) r4 = row;
    }
    if (nnew) {
#pragma unroll
        for (int t = 0; t < 5; ++t)
            if (W.gn(t) + ((cnt >> (6 * t)) & 63u) > (uint32_t)KB) return FC_SERIAL;
        if (used0 + nnew > W.cap) return FC_ROWS;  // (cap < 64, so no 6-bit count overflowed)
        // new slot i (first-appearance order) of table t takes the table's next slot and row
        // used0 + i; a reloaded slot keeps its slot and takes row used0 + i
        const uint32_t kv = ((uint32_t)l < nnew) ? F.key[l] : EMPTY;
        uint32_t seen = 0;
        for (uint32_t i = 0; i < nnew; ++i) {
            const uint32_t KF = rlu(kv, (int)i), K = KF & ~RELOAD;
            const int t = (int)(K >> 24);
            float* row = W.rows + (size_t)(used0 + i) * LY::LB;
            if (KF & RELOAD) {
                // the slot the decision loop found (the tables have not changed since)
                const uint64_t found = __ballot(l < 32 && W.gb(t) == (K & 0xFFFFFFu));
                const int slot = found ? __ffsll((unsigned long long)found) - 1 : 0;
                const uint32_t g = rlu(W.gr(t), slot);
                if (found && !(g & CACHED)) {
                    const float* src = (t == 0 ? psi_pool : bk_pool) + (size_t)g * LY::LB;
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) row[l + 64 * k] = src[l + 64 * k];
                    W.sr(t, CACHED | (used0 + i), l == slot);
                    if (l == 0) home[t * KB + slot] = g;
                }
            } else {
                const uint32_t slot = W.gn(t) + ((seen >> (6 * t)) & 63u);
                seen += 1u << (6 * t);
                W.sb(t, K & 0xFFFFFFu, (uint32_t)l == slot);
                W.sr(t, CACHED | (used0 + i), (uint32_t)l == slot);
#pragma unroll
                for (int k = 0; k < LY::PL; ++k) row[l + 64 * k] = 0.0f;
            }
            if (l == 0) {
                const uint32_t h = kc_slot(K);
                kc[h] = K;
                kc[NKC + h] = used0 + i;
            }
        }
#pragma unroll
        for (int t = 0; t < 5; ++t) W.sn(t, W.gn(t) + ((cnt >> (6 * t)) & 63u));
        W.used = used0 + nnew;
    }
    for (int rr = 0; rr < nrec; ++rr) {
        const uint32_t p = F.rec[l * FastLds<LY>::MR + rr];
        const uint32_t kidx = p >> 12;
        const uint32_t row = kidx == 0 ? r0 : kidx == 1 ? r1 : kidx == 2 ? r2 : kidx == 3 ? r3 : r4;
        F.rec[l * FastLds<LY>::MR + rr] = (uint16_t)(row * (uint32_t)LY::LB + (p & 0xFFFu));
    }
    return FC_OK;
}

// The claims of a batch (fast_claims_try), with one eviction of the cell's LDS rows when they
// run out: FC_OK, or FC_SERIAL for the serial path (a table would overflow, the batch alone needs
// more rows than the cell has, or a pool is full).
template <class LY>
__device__ inline int fast_claims(FastLds<LY>& F, WTabs<LY>& W, uint32_t* kc, uint32_t* home, Psi& out, Bkt& bk,
                                  uint32_t* flags) {
    const int l = lane_id();
    const int nkey = F.nkey[l], nrec = F.nrec[l];
    uint32_t k0 = EMPTY, k1 = EMPTY, k2 = EMPTY, k3 = EMPTY, k4 = EMPTY;
    if (nkey > 0) k0 = F.key[l * MAXKEY + 0];
    if (nkey > 1) k1 = F.key[l * MAXKEY + 1];
    if (nkey > 2) k2 = F.key[l * MAXKEY + 2];
    if (nkey > 3) k3 = F.key[l * MAXKEY + 3];
    if (nkey > 4) k4 = F.key[l * MAXKEY + 4];
    int fc = fast_claims_try<LY>(F, W, kc, home, out.pool, bk.pool, nkey, nrec, k0, k1, k2, k3, k4);
    if (fc == FC_ROWS && W.used > 0 && evict_rows<LY>(W, out, bk, home, kc, flags))
        fc = fast_claims_try<LY>(F, W, kc, home, out.pool, bk.pool, nkey, nrec, k0, k1, k2, k3, k4);
    return fc == FC_OK ? FC_OK : FC_SERIAL;
}

// Size of the per-workgroup owner table of the record rounds (hashed payload addresses), 2 KiB:
// addresses that share a slot only wait a round.
constexpr int NOWN_BITS = 7;  // 128 owner slots (512 B of LDS)
constexpr int NOWN = 1 << NOWN_BITS;

__device__ inline uint32_t own_slot(uint32_t addr) { return (addr * 2654435761u) >> (32 - NOWN_BITS); }

// The ordered adds, record-parallel: lane l holds component l's records (distinct
// addresses within a component).  Each round every pending record bids its index (component-major)
// for its address's owner slot; the lowest bid per slot adds its weight and leaves.  Records of
// one address therefore add in component order, one per round, as in the serial form; records
// of addresses that share a slot only wait.  own[] holds ~0u between rounds.
template <class LY>
__device__ inline void fast_rounds(FastLds<LY>& F, const WTabs<LY>& W, uint64_t m, uint32_t* own) {
    const int cnt = __popcll(m);
    const int l = lane_id();
    const int nr = (l < cnt) ? F.nrec[l] : 0;
    uint32_t pend = 0;
    for (int r = 0; r < nr; ++r)
        if (F.recw[l * FastLds<LY>::MR + r] > 0.0f) pend |= 1u << r;
    while (__ballot(pend != 0u)) {
        for (uint32_t q = pend; q; q &= q - 1) {
            const int r = __ffs(q) - 1;
            atomicMin(&own[own_slot(F.rec[l * FastLds<LY>::MR + r])], (uint32_t)(l * FastLds<LY>::MR + r));
        }
        uint32_t done = 0;
        for (uint32_t q = pend; q; q &= q - 1) {
            const int r = __ffs(q) - 1;
            const uint32_t ad = F.rec[l * FastLds<LY>::MR + r];
            if (own[own_slot(ad)] == (uint32_t)(l * FastLds<LY>::MR + r)) {
                __hip_atomic_fetch_add(W.rows + ad, F.recw[l * FastLds<LY>::MR + r], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
                done |= 1u << r;
            }
        }
        for (uint32_t q = pend; q; q &= q - 1) own[own_slot(F.rec[l * FastLds<LY>::MR + (__ffs(q) - 1)])] = ~0u;
        pend &= ~done;
    }
}

__device__ inline double rld(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ inline void lanes_to_accum(CellAccum& a, double accd, float accf) {
    a.edep = rld(accd, K_EDEP);
    a.E_cut = rld(accd, K_ECUT);
    a.E_nuc = rld(accd, K_ENUC);
    a.E_bnd = rld(accd, K_EBND);
    a.w_cut = rlf(accf, K_WCUT);
    a.w_nuc = rlf(accf, K_WNUC);
    a.w_bnd = rlf(accf, K_WBND);
    a.drop_w = rld(accd, K_DROPW);
    a.drop_E = rld(accd, K_DROPE);
    a.cl_win = rld(accd, K_CLWIN);
    a.cl_wout = rld(accd, K_CLWOUT);
    a.cl_Ein = rld(accd, K_CLEIN);
    a.cl_Eout = rld(accd, K_CLEOUT);
}

// Dynamic LDS of the transport: the energy edges and the range tables (4 * N_E + 1 floats), so
// that the binary searches and interpolations of every component read LDS.

// Deviation of a substep's learned widths from the plain physics of the same substep (E, ds in
// density rh and radiation length X0, mean energy loss mdE): the largest of |ln sigma_E|,
// |ln sigma_theta| differences and |a - a_plain| / (a_plain + 1e-4 / mm x rh ds), the quantities
// and the criterion of the table's node validation (k_table_plain).
__device__ inline float learned_deviation(const Tabs& T, float E, float ds, float rh, float X0, float mdE,
                                          float sE, float sth, float a) {
    const float sE_p = straggling(E, ds, rh);
    const float Em_p = fmaxf(E - 0.5f * mdE, T.ee[0]);
    const float sth_p = sqrtf_(fmaxf(scat_power(Em_p, ds, X0) * ds, 0.0f));
    const float a_p = nuc_sigma(E) * rh * ds;
    const float d0 = fabsf(logf_(fmaxf(sE, 1e-30f)) - logf_(fmaxf(sE_p, 1e-30f)));
    const float d1 = fabsf(logf_(fmaxf(sth, 1e-30f)) - logf_(fmaxf(sth_p, 1e-30f)));
    const float d2 = fabsf(a - a_p) / (a_p + 1e-4f * rh * ds);
    return fmaxf(d0, fmaxf(d1, d2));
}

// Per-bin physics of one component (the per-bin algorithm of k2_coarsetransport.cu:76-704 and
// k3_finetransport.cu:94-760) plus the table-independent parts of its apply: one lane, one
// (slot, local bin) of the listed cell.
struct CellGeo {
    int lc, iz, ix, Nx, NT, NE;
    float dx, dz, hdx, hdz, tmin, dth, rh, X0, A, B, sx0;
    float R_on;  // water range at the fine-transport threshold energy (a launch constant)
    uint32_t latlisted;
    int monitor;  // the learned physics of this cell's substeps is compared with the plain physics
};

// Learned widths of a component evaluated by the network itself (learned-physics mode 2).
struct Learned {
    int use;  // the substep takes these widths (inputs inside the training range)
    float sE, sth, a;
};

// Mode 2: the substep inputs of every lane's fine-transport component (the energy at its bin
// centre and the step ds as transport_component computes them, dR = rho ds, rho, X0 ratio) go
// through the network on the matrix cores, 64 components per wavefront.  Every lane of the
// wavefront must call it.
template <class LY>
__device__ Learned learned_inline(const Tabs& T, const Tp& P, const CellGeo& G, bool live, float w_in, int lidx,
                                  uint32_t bT, uint32_t bEb) {
    float x[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    bool use = false;
    if (live && !(w_in < 1e-15f)) {
        const int tbin = max(0, min((int)bT * LY::TL + LY::lt(lidx), G.NT - 1));
        const int ebin = max(0, min((int)bEb * LY::EL + LY::le(lidx), G.NE - 1));
        const float E = ecenter(T.ee, ebin);
        if (!(E <= 0.1f)) {
            const float theta = T.te[tbin] + 0.5f * G.dth;
            const float xc = G.dx * ((float)(2 * LY::lx(lidx) + 1 - LY::XS) / (float)(2 * LY::XS));
            const float zc = G.dz * ((float)(2 * LY::lz(lidx) + 1 - LY::ZS) / (float)(2 * LY::ZS));
            float mu, eta;
            sincosf_(theta, eta, mu);
            const float s1 = (mu > 0) ? (G.hdz - zc) / mu : 1e30f;
            const float s2 = (mu < 0) ? (-G.hdz - zc) / mu : 1e30f;
            const float s3 = (eta > 0) ? (G.hdx - xc) / eta : 1e30f;
            const float s4 = (eta < 0) ? (-G.hdx - xc) / eta : 1e30f;
            const float sb = fmaxf(fminf(fminf(s1, s2), fminf(s3, s4)), 0.0f);
            const float ds = fminf(max_step(T, E, G.rh, P.rule, ebin), sb * 1.001f);
            use = T.rowmat && T.rowmat[G.iz] >= 0 && ds > 0.0f && ds / G.X0 >= 1e-6f &&
                  mlp::encode_input(E, G.rh * ds, G.rh, T.x0r[G.iz], x);
        }
    }
    float lsE, lsT, a;
    mlp::mlp_wave_forward(T.mlp_wb, T.mlp_p, T.mlp_dict, x, lsE, lsT, a);
    Learned o;
    o.use = use ? 1 : 0;
    o.sE = expf_(lsE);
    o.sth = expf_(lsT);
    o.a = fmaxf(a, 0.0f);
    return o;
}

// Mode-2 widths of the wavefront's components, or none (the table or the plain physics).  INFER
// selects the kernel instances that carry the network: its inference needs most of the register
// file, so the other instances do not contain it.
template <class LY, bool FINE, bool INFER>
__device__ inline Learned learned_for(const Tabs& T, const Tp& P, const CellGeo& G, bool live, float w_in, int lidx,
                                      uint32_t bT, uint32_t bEb) {
    if constexpr (FINE && INFER) {
        if (P.surr == 2) return learned_inline<LY>(T, P, G, live, w_in, lidx, bT, bEb);
    }
    Learned o;
    o.use = 0;
    o.sE = o.sth = o.a = 0.0f;
    return o;
}

template <class LY, bool FINE>
__device__ inline Comp transport_component(const Tabs& T, const Tp& P, const CellGeo& G, bool live, float w_in,
                                           int lidx_in, uint32_t slot, uint32_t bT, uint32_t bEb,
                                           const Learned& Lx SM2D_SUBP) {
    SM2D_SUB_START();
    const int lc = G.lc, iz = G.iz, ix = G.ix, Nx = G.Nx, NT = G.NT, NE = G.NE;
    const float dx = G.dx, dz = G.dz, hdx = G.hdx, hdz = G.hdz, tmin = G.tmin, dth = G.dth;
    const float rh = G.rh, X0 = G.X0, A = G.A, B = G.B, sx0 = G.sx0;
    const uint32_t latlisted = G.latlisted;
    Comp r;
    int eg = -1;  // energy bin of the component (a starting guess for the bin searches)
    r.kind = 2;
    r.face = -1;
    r.fmask = 0;
    r.surr = 0;
    r.sdev = 0.0f;
    r.lidx = lidx_in;
    r.w = live ? w_in : 0.0f;
    r.edep = r.w_rem = r.E_rem = r.w_new = r.E_new = 0.0f;
    r.th_new = r.x_new = r.z_new = 0.0f;
    r.wcf = r.wl = r.wr = 0.0f;
#pragma unroll
    for (int q = 0; q < 11; ++q) r.sw[q] = 0.0f;
    if (live) {
        const int lidx = r.lidx;
        const float weight = w_in;
        int tbin = (int)bT * LY::TL + LY::lt(lidx);
        int ebin = (int)bEb * LY::EL + LY::le(lidx);
        if (FINE) {
            tbin = max(0, min(tbin, NT - 1));
            ebin = max(0, min(ebin, NE - 1));
        }
        const float E = ecenter(T.ee, ebin);
        eg = ebin;
        r.E = E;
        const bool pruned = FINE ? (weight < 1e-15f) : (weight < 1e-12f);
        if (pruned) {
            r.kind = 0;
        } else if (E <= 0.1f) {
            r.kind = 1;
        } else {
            const float theta = FINE ? T.te[tbin] + 0.5f * dth : tmin + (tbin + 0.5f) * dth;
            const float xc = dx * ((float)(2 * LY::lx(lidx) + 1 - LY::XS) / (float)(2 * LY::XS));
            const float zc = dz * ((float)(2 * LY::lz(lidx) + 1 - LY::ZS) / (float)(2 * LY::ZS));
            float mu, eta;
            sincosf_(theta, eta, mu);
            SM2D_SUB(0);
            float ds, dE, E_new, edep, w_new, w_rem, E_rem, sx, sxt, sth;
            unsigned pseed;
            if (FINE) {
                float s1 = (mu > 0) ? (hdz - zc) / mu : 1e30f;
                float s2 = (mu < 0) ? (-hdz - zc) / mu : 1e30f;
                float s3 = (eta > 0) ? (hdx - xc) / eta : 1e30f;
                float s4 = (eta < 0) ? (-hdx - xc) / eta : 1e30f;
                float sb = fmaxf(fminf(fminf(s1, s2), fminf(s3, s4)), 0.0f);
                ds = fminf(max_step(T, E, rh, P.rule, ebin), sb * 1.001f);
                float mdE = deposition(T, E, ds, rh, ebin);
                float s_sE = 0.0f, s_sth = 0.0f, s_a = 0.0f;
                bool use_s;
                if (P.surr == 2) {  // the network, evaluated for the wavefront (learned_inline)
                    use_s = Lx.use != 0;
                    s_sE = Lx.sE;
                    s_sth = Lx.sth;
                    s_a = Lx.a;
                } else {
                    use_s = P.surr && ds > 0.0f && ds / X0 >= 1e-6f &&
                            surrogate_lookup(T, ebin, iz, rh * ds, s_sE, s_sth, s_a);
                }
                r.surr = use_s ? 1 : 0;
                if (P.strag) {
                    unsigned sd = (unsigned)((lc * 7 + (int)slot * 13 + lidx * 17) ^ 0x5DEECE66DL);
                    float sE = use_s ? s_sE : straggling(E, ds, rh);
                    dE = (sE <= 0.0f) ? mdE : fmaxf(mdE + gauss(sd) * sE, 0.0f);
                } else {
                    dE = mdE;
                }
                dE = fminf(dE, E);
                E_new = E - dE;
                edep = dE * weight;
                if (P.nuc) {
                    float surv = use_s ? expf_(-s_a) : expf_(-(nuc_sigma(E) * rh) * ds);
                    w_new = weight * surv;
                    w_rem = weight - w_new;
                    E_rem = w_rem * E;
                } else {
                    w_new = weight;
                    w_rem = 0.0f;
                    E_rem = 0.0f;
                }
                float smm = fmaxf(ds, 0.0f);
                float Em = fmaxf(E - 0.5f * dE, T.ee[0]);
                if (use_s && G.monitor) r.sdev = learned_deviation(T, E, ds, rh, X0, mdE, s_sE, s_sth, s_a);
                float Tp_ = use_s ? (s_sth * s_sth) / smm : scat_power(Em, smm, X0);
                sth = use_s ? s_sth : sqrtf_(fmaxf(Tp_ * smm, 0.0f));
                float dC = 2.0f * B * smm + A * smm * smm + Tp_ * smm * smm * smm / 3.0f;
                dC = fmaxf(dC, 0.0f);
                sx = fmaxf(sqrtf_(dC), 0.01f);
                sxt = 0.001f;
                pseed = (unsigned)fminf((float)iz * dz * 1000.0f, 4294967295.0f);
            } else {
                float cstep = fminf(P.step_coarse, fminf(dx, dz));
                float crs = cstep;
                bool split = false;
                if (E > P.E_on + 1e-6f) {
                    float Rc = lookup_R(T, E, ebin);
                    float Rf = G.R_on;
                    float rtf = (Rc - Rf) / rh;
                    if (rtf > 0.0f && rtf < crs) {
                        crs = rtf;
                        cstep = crs;
                        split = true;
                    }
                }
                dE = fminf(deposition(T, E, crs, rh, ebin), E);
                E_new = E - dE;
                edep = dE * weight;
                float surv = expf_(-(nuc_sigma(E) * rh) * crs);
                w_new = weight * surv;
                w_rem = weight - w_new;
                E_rem = w_rem * E;
                if (split && E_new > P.E_on) {
                    float ex = E_new - P.E_on;
                    if (ex > 0.0f && w_new > 0.0f) edep += ex * w_new;
                    E_new = P.E_on;
                }
                float ps = fmaxf(0.0f, (float)iz * dz + (zc + 0.5f * dz));
                float smm = fmaxf(crs, 0.0f);
                float st0 = (ps > 0.0f) ? highland(E, ps, X0) : 0.0f;
                float Ao = st0 * st0;
                float Bo = 0.5f * Ao * ps;
                float Co = sx0 * sx0 + (Ao * ps * ps) / 3.0f;
                float Em = fmaxf(E - 0.5f * dE, T.ee[0]);
                float Tp_ = scat_power(Em, smm, X0);
                float An = fmaxf(Ao + Tp_ * smm, 0.0f);
                float Cn = fmaxf(Co + 2.0f * Bo * smm + Ao * smm * smm +
                                     (1.0f / 3.0f) * Tp_ * smm * smm * smm,
                                 0.0f);
                sth = sqrtf_(fmaxf(An - Ao, 0.0f));
                float dC = fmaxf(Cn - Co, 0.0f);
                float unres = fminf(0.35f, 0.04f + 0.003f * ps);
                sx = fmaxf(sqrtf_(fmaxf(dC + unres * Co, 0.0f)), 0.01f);
                sxt = fmaxf(sinf_(fminf(sth, 1.57f)) * smm / 1.7320508f, 0.01f);
                pseed = (unsigned)fminf(ps * 1000.0f, 4294967295.0f);
                ds = cstep;
            }
            SM2D_SUB(1);
            unsigned ts = ((unsigned)(lc + 1) * 2654435761u) ^
                          ((unsigned)(slot + 1) * 2246822519u) ^
                          ((unsigned)(lidx + 1) * 3266489917u) ^ (unsigned)(ebin + 17) ^
                          (pseed * 747796405u);
            float tsc = (sth <= 0.0f) ? 0.0f : gauss(ts) * sth;
            float thn = fmaxf(T.te[0], fminf(theta + tsc, T.te[NT]));
            float mun, etn;
            sincosf_(thn, etn, mun);
            float xn = xc + etn * ds, zn = zc + mun * ds;
            int face = -1;
            if (zn >= hdz - 0.001f) face = 0;
            else if (zn <= -hdz + 0.001f) face = 1;
            else if (xn >= hdx - 0.001f) face = 2;
            else if (xn <= -hdx + 0.001f) face = 3;
            xn = fmaxf(-hdx, fminf(xn, hdx));
            zn = fmaxf(-hdz, fminf(zn, hdz));
            r.edep = edep;
            r.w_rem = w_rem;
            r.E_rem = E_rem;
            r.w_new = w_new;
            r.E_new = E_new;
            r.face = face;
            r.th_new = thn;
            r.x_new = xn;
            r.z_new = zn;
            SM2D_SUB(2);
            if (E_new > 0.1f && face <= 0) {
                // Both spreads need Gaussian CDF values at up to 12 points; they are evaluated
                // in one uniform loop so that a wave holding both kinds of component runs the
                // CDF once per point, not once per branch.
                // face 0: device_emit_lateral_spread weights (device_bucket.cuh:880-916), 11
                // cells, points xmin + k dx (k = 0..11), width sxt.
                // inside: device_gaussian_spread_weights_subcell + in-cell fraction and tails
                // (k3_finetransport.cu:585-653), points -dx/2 + k dx/XS (k = 0..XS), width
                // s, then +-dx/2 with width sx.
                const bool f0 = face == 0;
                const float s0w = fmaxf(sxt, 1e-6f);
                const float s1w = fmaxf(FINE ? sxt : sx, 1e-6f);
                const float dxs = dx / LY::XS;
                const float xa = f0 ? xn - 0.5f * 11.0f * dx : -dx * 0.5f;
                float g[12];
#pragma unroll
                for (int q = 0; q < 12; ++q) {
                    float x, sq;
                    if (f0) {
                        x = q == 0 ? xa : xa + q * dx;
                        sq = s0w;
                    } else if (q <= LY::XS) {
                        x = q == 0 ? xa : xa + q * dxs;
                        sq = s1w;
                    } else if (q == LY::XS + 1) {
                        x = hdx;
                        sq = sx;
                    } else if (q == LY::XS + 2) {
                        x = -hdx;
                        sq = sx;
                    } else {
                        x = xn;
                        sq = 1.0f;
                    }
                    g[q] = gcdf(x, xn, sq);
                }
                if (f0) {
                    float sum = 0.0f;
#pragma unroll
                    for (int q = 0; q < 11; ++q) r.sw[q] = g[q + 1] - g[q];
#pragma unroll
                    for (int q = 0; q < 11; ++q) sum += r.sw[q];
                    if (sum > 1e-10f) {
#pragma unroll
                        for (int q = 0; q < 11; ++q) r.sw[q] /= sum;
                    } else {
#pragma unroll
                        for (int q = 0; q < 11; ++q) r.sw[q] = (q == 5) ? 1.0f : 0.0f;
                    }
                    sum = 0.0f;
#pragma unroll
                    for (int q = 0; q < 11; ++q) {
                        r.sw[q] = fmaxf(0.0f, r.sw[q]);
                        sum += r.sw[q];
                    }
                    if (sum < 1e-12f) {
#pragma unroll
                        for (int q = 0; q < 11; ++q) r.sw[q] = (q == 5) ? 1.0f : 0.0f;
                    } else {
#pragma unroll
                        for (int q = 0; q < 11; ++q) r.sw[q] /= sum;
                    }
                } else {
                    float sum = 0.0f;
                    float w8[LY::XS];
#pragma unroll
                    for (int q = 0; q < LY::XS; ++q) w8[q] = g[q + 1] - g[q];
#pragma unroll
                    for (int q = 0; q < LY::XS; ++q) sum += w8[q];
                    if (sum > 1e-10f) {
#pragma unroll
                        for (int q = 0; q < LY::XS; ++q) w8[q] /= sum;
                    } else {
#pragma unroll
                        for (int q = 0; q < LY::XS; ++q) w8[q] = (q == LY::XS / 2) ? 1.0f : 0.0f;
                    }
                    const float gp = g[LY::XS + 1], gm = g[LY::XS + 2];  // CDF at +dx/2 and -dx/2
                    float win = gp - gm;
                    float wcf = fminf(1.0f, fmaxf(0.0f, win));
                    float wss = 0.0f;
#pragma unroll
                    for (int q = 0; q < LY::XS; ++q) {
                        w8[q] = fmaxf(0.0f, w8[q]);
                        wss += w8[q];
                    }
                    if (wss < 1e-12f) {
#pragma unroll
                        for (int q = 0; q < LY::XS; ++q) w8[q] = (q == LY::XS / 2) ? 1.0f : 0.0f;
                        wss = 1.0f;
                    }
#pragma unroll
                    for (int q = 0; q < LY::XS; ++q) r.sw[q] = (w8[q] / wss) * wcf;
                    float wl = 0.0f, wr = 0.0f;
                    if (wcf < 1.0f - 1e-6f) {
                        float wlr = fmaxf(0.0f, gm);
                        float wrr = fmaxf(0.0f, 1.0f - gp);
                        float wtt = fmaxf(0.0f, 1.0f - wcf);
                        float wts = wlr + wrr;
                        if (wtt > 0.0f && wts > 1e-12f) {
                            float sc = wtt / wts;
                            wl = wlr * sc;
                            wr = wrr * sc;
                        }
                    }
                    r.wcf = wcf;
                    r.wl = wl;
                    r.wr = wr;
                }
            }
        }
    }
    SM2D_SUB(3);
    // ------- emission binning and the table-independent parts of the apply -------
    r.bn = Bins{0, 0, 0, 0.0f, 0.0f, 0.0f};
    r.i0 = r.i1 = r.i2 = 0;
    r.g0 = r.g1 = r.g2 = 0.0f;
    if (live && r.kind == 2 && r.E_new > 0.1f) {
        r.bn = emission_bins(T, r.th_new, r.E_new, eg);
        const float w_new = r.w_new;
        if (r.face >= 0) {
            r.g0 = ecenter(T.ee, find_bin(T.ee, NE, r.E_new, eg));
            if (r.face == 0) {
                float drop = 0.0f, lat_bnd = 0.0f;
                uint32_t fm = 0;
#pragma unroll
                for (int q = 0; q < 11; ++q) {
                    const float wf = r.sw[q];
                    if (wf <= 0.0f) continue;
                    const int it = ix + q - 5;
                    if (it < 0 || it >= Nx) {
                        lat_bnd += w_new * wf;
                        continue;
                    }
                    if (q == 5) {
                        drop += 0.0f;  // the centre emission's drop, zero unless a pool fills
                    } else if ((latlisted >> q) & 1u) {
                        // another listed cell's +z bucket: emitted after every cell of the
                        // phase has made its own emissions (apply_foreign)
                        fm |= 1u << q;
                    } else {
                        // a cell not listed in this phase has no bucket: a drop
                        // (device_bucket.cuh:953-963)
                        drop += w_new * wf;
                    }
                }
                r.fmask = fm;
                r.g1 = drop;
                r.g2 = lat_bnd;
            } else {
                const int face = r.face;
                const float xon = (face == 2) ? -dx * 0.5f : (face == 3 ? dx * 0.5f : r.x_new);
                r.i0 = min(max((int)(((xon / dx) + 0.5f) * LY::XS), 0), LY::XS - 1);
                if (face == 1) r.i1 = LY::ZS - 1;
                else r.i1 = min(max((int)(((r.z_new / dz) + 0.5f) * LY::ZS), 0), LY::ZS - 1);
            }
        } else {
            int tbn = (int)((r.th_new - tmin) / dth);
            tbn = (int)fmaxf(0, fminf(tbn, NT - 1));
            const int ebn = find_bin(T.ee, NE, r.E_new, eg);
            r.i0 = tbn;
            r.i1 = ebn;
            r.i2 = min(max((int)(((r.z_new / dz) + 0.5f) * LY::ZS), 0), LY::ZS - 1);
            r.g0 = ecenter(T.ee, ebn);
        }
    }
    SM2D_SUB(4);
    return r;
}

// The serial form of the ordered apply (components in bin order, one at a time): used for a
// batch whose claims need the closest-block fallback or a global payload.  Component j's
// accumulator contributions are formed in `a` (all lanes, from zero) in the serial order.
template <class LY, bool FINE>
__device__ inline void serial_component(const Comp& r, int j, WTabs<LY>& W, CellAccum& a, const Tabs& T, Psi& out,
                                        Bkt& bk, uint32_t* flags, const CellGeo& G) {
    const int l = lane_id();
    const int ix = G.ix, iz = G.iz, Nx = G.Nx, NT = G.NT;
    const float tmin = G.tmin, dth = G.dth;
    (void)NT;
    (void)tmin;
    (void)dth;
    const int kind = rli(r.kind, j);
    const float w = rlf(r.w, j), E = rlf(r.E, j);
    a.cl_win += (double)w;
    a.cl_Ein += (double)w * (double)E;
    if (kind == 0) {
        a.drop_w += (double)w;
        a.drop_E += (double)(E * w);
        a.cl_wout += (double)w;
        a.cl_Eout += (double)(E * w);
        return;
    }
    if (kind == 1) {
        a.w_cut += w;
        a.E_cut += E * w;
        a.cl_wout += (double)w;
        a.cl_Eout += (double)(E * w);
        return;
    }
    const float edep = rlf(r.edep, j), w_rem = rlf(r.w_rem, j), E_rem = rlf(r.E_rem, j);
    const float w_new = rlf(r.w_new, j), E_new = rlf(r.E_new, j);
    const int face = rli(r.face, j);
    a.edep += edep;
    a.w_nuc += w_rem;
    a.E_nuc += E_rem;
    a.cl_wout += (double)w_rem;
    a.cl_Eout += (double)edep + (double)E_rem;
    if (E_new <= 0.1f) {
        a.w_cut += w_new;
        a.E_cut += E_new * w_new;
        a.cl_wout += (double)w_new;
        a.cl_Eout += (double)(E_new * w_new);
        return;
    }
    Bins bn;
    bn.bc = rli(r.bn.bc, j);
    bn.tb1 = rli(r.bn.tb1, j);
    bn.eb = rli(r.bn.eb, j);
    bn.tw0 = rlf(r.bn.tw0, j);
    bn.tw1 = rlf(r.bn.tw1, j);
    bn.Ecen = rlf(r.bn.Ecen, j);
    if (face >= 0) {
        bool nb = (face == 0) ? (iz + 1 < T.Nz) : (face == 1) ? (iz > 0)
                : (face == 2) ? (ix + 1 < Nx) : (ix > 0);
        if (!nb) {
            a.w_bnd += w_new;
            a.E_bnd += E_new * w_new;
            a.cl_wout += (double)w_new;
            a.cl_Eout += (double)(E_new * w_new);
            return;
        }
        const float E_emit = rlf(r.g0, j);
        float drop = 0.0f;
        if (face == 0) {
            const float wf5 = rlf(r.sw[5], j);
            const float lat_bnd = rlf(r.g2, j);
            const uint32_t fmj = rlu(r.fmask, j);
            float d5 = 0.0f;
            if (wf5 > 0.0f) d5 = emit<LY>(W, a, T, bk, flags, 0, bn, w_new * wf5, LY::XS / 2, 0);
#pragma unroll
            for (int q = 0; q < 11; ++q) {  // lateral emissions into other listed cells
                if (!((fmj >> q) & 1u)) continue;
                const float wl = w_new * rlf(r.sw[q], j);
                if (bn.tw0 > 0.0f) {
                    const float we = wl * bn.tw0;
                    a.cl_wout += (double)we;
                    a.cl_Eout += (double)we * (double)bn.Ecen;
                }
                if (bn.tw1 > 0.0f) {
                    const float we = wl * bn.tw1;
                    a.cl_wout += (double)we;
                    a.cl_Eout += (double)we * (double)bn.Ecen;
                }
            }
            if (d5 == 0.0f) {
                drop = rlf(r.g1, j);
            } else {  // the lateral sum with the centre's drop in its place (q = 5)
#pragma unroll
                for (int q = 0; q < 11; ++q) {
                    const float wf = rlf(r.sw[q], j);
                    if (wf <= 0.0f || ((fmj >> q) & 1u)) continue;
                    const int it = ix + q - 5;
                    if (it < 0 || it >= Nx) continue;
                    drop += (q == 5) ? d5 : w_new * wf;
                }
            }
            if (lat_bnd > 0.0f) {
                a.w_bnd += lat_bnd;
                a.E_bnd += E_new * lat_bnd;
                a.cl_wout += (double)lat_bnd;
                a.cl_Eout += (double)(E_new * lat_bnd);
            }
            if (drop > 0.0f) {
                a.drop_w += (double)drop;
                a.drop_E += (double)(E_new * drop);
                a.cl_wout += (double)drop;
                a.cl_Eout += (double)(E_new * drop);
            }
            if (FINE) {
                float tw = fmaxf(0.0f, w_new - lat_bnd - drop);
                if (tw > 0.0f) {
                    a.edep += (double)((E_new - E_emit) * tw);
                    a.cl_Eout += (double)((E_new - E_emit) * tw);
                }
            }
        } else {
            const int xsn = rli(r.i0, j), zsn = rli(r.i1, j);
            drop = emit<LY>(W, a, T, bk, flags, face, bn, w_new, xsn, zsn);
            if (drop > 0.0f) {
                a.drop_w += (double)drop;
                a.drop_E += (double)(E_new * drop);
                a.cl_wout += (double)drop;
                a.cl_Eout += (double)(E_new * drop);
            }
            if (FINE) {
                float tw = fmaxf(0.0f, w_new - drop);
                if (tw > 0.0f) {
                    a.edep += (double)((E_new - E_emit) * tw);
                    a.cl_Eout += (double)((E_new - E_emit) * tw);
                }
            }
        }
        return;
    }
    // ---- remain in the cell ----
    const int tbn = rli(r.i0, j), ebn = rli(r.i1, j), zst = rli(r.i2, j);
    const uint32_t bidn = ((uint32_t)(ebn / LY::EL) << 12) | ((uint32_t)(tbn / LY::TL) & 0xFFF);
    uint32_t sel;
    const int os = W.claim(0, bidn, sel, out.pool, out.pcount, out.pcap, flags, F_PSI_FULL);
    if (os < 0) {
        a.drop_w += (double)w_new;
        a.drop_E += (double)(E_new * w_new);
        a.cl_wout += (double)w_new;
        a.cl_Eout += (double)(E_new * w_new);
        return;
    }
    int tbe = tbn, ebe = ebn;
    float E_emit = rlf(r.g0, j);
    if (sel != bidn) {
        int bte = (int)(sel & 0xFFF), bee = (int)((sel >> 12) & 0xFFF);
        tbe = max(bte * LY::TL, min(tbe, bte * LY::TL + LY::TL - 1));
        ebe = max(bee * LY::EL, min(ebe, bee * LY::EL + LY::EL - 1));
        E_emit = T.ee[ebe] + 0.5f * (T.ee[ebe + 1] - T.ee[ebe]);
    }
    const uint32_t oref = rlu(W.r0, os);
    float mywf = 0.0f;
#pragma unroll
    for (int q = 0; q < LY::XS; ++q) {
        float wf = rlf(r.sw[q], j);
        if (l == q) mywf = wf;
        if (wf > 0.0f) {
            float wsp = w_new * wf;
            a.cl_wout += (double)wsp;
            a.cl_Eout += (double)wsp * (double)E_emit;
        }
    }
    if (l < LY::XS && mywf > 0.0f)
        W.add(oref, LY::enc(tbe % LY::TL, ebe % LY::EL, l, zst), w_new * mywf, out.pool);
    const float wcf = rlf(r.wcf, j), wl = rlf(r.wl, j), wr = rlf(r.wr, j);
    if (wcf < 1.0f - 1e-6f) {
        const int zsb = zst;
        if (wl > 0.0f) {
            float wsl = w_new * wl;
            if (ix > 0) {
                float d = emit<LY>(W, a, T, bk, flags, 3, bn, wsl, LY::XS - 1, zsb);
                if (d > 0.0f) {
                    a.drop_w += (double)d;
                    a.drop_E += (double)(E_new * d);
                    a.cl_wout += (double)d;
                    a.cl_Eout += (double)(E_new * d);
                }
            } else {
                a.w_bnd += wsl;
                a.E_bnd += E_new * wsl;
                a.cl_wout += (double)wsl;
                a.cl_Eout += (double)(E_new * wsl);
            }
        }
        if (wr > 0.0f) {
            float wsr = w_new * wr;
            if (ix < Nx - 1) {
                float d = emit<LY>(W, a, T, bk, flags, 2, bn, wsr, 0, zsb);
                if (d > 0.0f) {
                    a.drop_w += (double)d;
                    a.drop_E += (double)(E_new * d);
                    a.cl_wout += (double)d;
                    a.cl_Eout += (double)(E_new * d);
                }
            } else {
                a.w_bnd += wsr;
                a.E_bnd += E_new * wsr;
                a.cl_wout += (double)wsr;
                a.cl_Eout += (double)(E_new * wsr);
            }
        }
    }
    if (FINE) {
        float rf = wcf;
        if (ix > 0) rf += wl;
        if (ix < Nx - 1) rf += wr;
        rf = fminf(1.0f, fmaxf(0.0f, rf));
        float rw = w_new * rf;
        if (rw > 0.0f) {
            a.edep += (double)((E_new - E_emit) * rw);
            a.cl_Eout += (double)((E_new - E_emit) * rw);
        }
    }
}

// The batch's components in bin order; lane j of `mine` receives component j's contributions.
template <class LY, bool FINE>
__device__ inline void serial_apply(const Comp& r, uint64_t m, WTabs<LY>& W, CellAccum& mine, uint32_t& comps,
                                    const Tabs& T, Psi& out, Bkt& bk, uint32_t* flags, const CellGeo& G) {
    const int l = lane_id();
    while (m) {
        const int j = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        CellAccum a = {};
        serial_component<LY, FINE>(r, j, W, a, T, out, bk, flags, G);
        ++comps;
        if (l == j) mine = a;
    }
}

__device__ inline void set_cell_geo(CellGeo& G, const Tabs& T, int cell, const uint32_t* ltag, uint32_t tag) {
    const int l = lane_id();
    const int sp = cell / T.Ns;
    G.lc = cell - sp * T.Ns;  // cell within the spot
    G.iz = G.lc / G.Nx;
    G.ix = G.lc - G.iz * G.Nx;
    G.rh = T.rho[G.iz];
    G.X0 = X0W * T.x0r[G.iz];
    G.A = T.feA[sp * T.Nz + G.iz];
    G.B = T.feB[sp * T.Nz + G.iz];
    G.sx0 = T.sx0[sp];
    G.monitor = (G.lc % SURR_MONITOR_EVERY) == 0;
    const int off = l - 5, it = G.ix + off;
    const bool q = l < 11 && off != 0 && it >= 0 && it < G.Nx && ltag[cell + off] == tag;
    G.latlisted = (uint32_t)__ballot(q);
}

// The lateral emissions of a phase into other listed cells' +z buckets (device_emit_lateral_spread,
// device_bucket.cuh:952-983), applied by one wavefront once every cell of the phase has made its
// own emissions.  For each marked target bucket, the sources S (the listed cells up to five to
// either side, in ascending cell order) are transported again, component by component in their
// serial order, with the same function and inputs (so the same results), and each component's
// emission toward the target is claimed and added in that order: source-cell order, component
// order, theta part.  This is one of the orders the reference's device atomics can produce: each
// bucket location orders its relaxed atomics independently, so the bucket's own cell may come
// first.  The weights were booked to their sources when emitted.
template <class LY, bool FINE, bool INFER>
__device__ void apply_foreign(const Tabs& T, const Tp& P, CellGeo G, const Psi& in, Bkt& bk,
                              const uint32_t* __restrict__ list, const uint32_t* __restrict__ ltag, uint32_t tag,
                              uint32_t nf, uint32_t* flags) {
    const int l = lane_id();
    for (uint32_t k = 0; k < nf; ++k) {
        const uint32_t pT = bk.flist[k];
        const int tc = (int)list[pT];
        const int ixT = (tc % T.Ns) % T.Nx;
        const uint32_t base = (pT * 4 + 0) * KB;  // the target's +z bucket (face 0)
        uint32_t n = bk.n[pT * 4];
        uint32_t mybid = (l < 32 && (uint32_t)l < n) ? bk.bid[base + l] : EMPTY;
        uint32_t myref = (l < 32 && (uint32_t)l < n) ? bk.sref[base + l] : 0u;
        for (int off = 5; off >= -5; --off) {  // source S = target - off, ascending
            if (off == 0) continue;
            const int ixS = ixT - off;
            if (ixS < 0 || ixS >= T.Nx) continue;
            const int S = tc - off;
            if (ltag[S] != tag) continue;
            set_cell_geo(G, T, S, ltag, tag);
            if (G.iz + 1 >= T.Nz) continue;  // +z exits at the last row leave the grid
            const int q = off + 5;           // the target's index in the source's lateral spread
            const int xs = off > 0 ? 0 : LY::XS - 1;
            const uint32_t ns = in.ns[S];
            for (uint32_t sl = 0; sl < ns; ++sl) {
                const uint32_t bid = in.bid[(size_t)S * KB + sl];
                if (bid == EMPTY) continue;
                const float* vin = in.pool + (size_t)in.sref[(size_t)S * KB + sl] * LY::LB;
                for (int kk = 0; kk < LY::PL; ++kk) {  // local bins ascending
                    const float w = vin[l + 64 * kk];
                    const bool live = w > 0.0f;
                    if (!__ballot(live)) continue;
                    const Learned Lx = learned_for<LY, FINE, INFER>(T, P, G, live, live ? w : 0.0f, l + 64 * kk,
                                                             bid & 0xFFF, (bid >> 12) & 0xFFF);
                    const Comp r = transport_component<LY, FINE>(T, P, G, live, live ? w : 0.0f, l + 64 * kk, sl,
                                                                 bid & 0xFFF, (bid >> 12) & 0xFFF,
                                                                 Lx SM2D_SUBA(nullptr));
                    float wq = 0.0f;
#pragma unroll
                    for (int u = 0; u < 11; ++u)
                        if (u == q) wq = r.sw[u];
                    const bool em = live && r.kind == 2 && r.E_new > 0.1f && r.face == 0 && ((r.fmask >> q) & 1u);
                    for (uint64_t m = __ballot(em); m; m &= m - 1) {
                        const int j = __ffsll((unsigned long long)m) - 1;
                        const float wl = rlf(r.w_new, j) * rlf(wq, j);
                        const int bc = rli(r.bn.bc, j), tb1 = rli(r.bn.tb1, j), eb = rli(r.bn.eb, j);
                        const float tw0 = rlf(r.bn.tw0, j), tw1 = rlf(r.bn.tw1, j);
                        const int el = eb % LY::EL;
                        const uint32_t bE = (uint32_t)(eb / LY::EL);
#pragma unroll
                        for (int ti = 0; ti < 2; ++ti) {
                            const float tw = ti == 0 ? tw0 : tw1;
                            if (tw <= 0.0f) continue;
                            const float we = wl * tw;
                            const int t = ti == 0 ? bc : tb1;
                            const uint32_t want = ((uint32_t)(t / LY::TL) & 0xFFF) | ((bE & 0xFFF) << 12);
                            const int lidx = LY::enc(t % LY::TL, el, xs, 0);
                            uint32_t sel;
                            const int s2 = table_find_or_claim(mybid, myref, n, want, sel, bk.pool, bk.pcount, bk.pcap,
                                                               bk.bid + base, bk.sref + base, LY::LB, flags, F_BKT_FULL);
                            if (s2 < 0) continue;  // the bucket pool is full: the run ends with its flag
                            const int l2 = (sel == want) ? lidx : remap<LY>(lidx, want, sel);
                            const uint32_t ref = rlu(myref, s2);
                            if (l == 0) bk.pool[(size_t)ref * LY::LB + l2] += we;
                        }
                    }
                }
            }
        }
        if (l == 0) bk.n[pT * 4] = n;
    }
}

// Item queue of a cell: its non-zero (slot, local bin) components in the serial order
// (slots ascending, local bins ascending), processed in batches of 64 by the four waves.
// TW waves work on one listed cell: they evaluate the physics of TW batches at once, wave 0
// applies them in order.  TW is a launch choice (1, 2 or 4) made by the runtime's tuner; all
// choices give identical results.
constexpr int QCAP = 512;  // non-zero components queued per round (at least one slot's worth)

// Dynamic LDS of a transport launch: the energy tables, for the 4-wave shape only.
inline size_t transport_dyn_lds(int NE, int TW) { return TW == 4 ? sizeof(float) * (4 * (size_t)NE + 1) : 0; }

template <class LY, bool FINE, int TW, bool INFER = false>
__global__ __launch_bounds__(64 * TW) __attribute__((amdgpu_waves_per_eu(2))) void k_transport(Tabs Tg, Tp P, Psi in, Psi out, Bkt bk,
                                                   const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ nlist,
                                                   const uint32_t* __restrict__ ltag,
                                                   const uint32_t* __restrict__ lpos, uint32_t* ctl,
                                                   Tally tal, double* acc, uint32_t* flags) {
    if (ctl[C_DONE] || blockIdx.x >= *nlist) return;  // no cell for this workgroup
    // LDS payload rows: 4-wave workgroups (few listed cells, one workgroup per CU) keep the full
    // set (24 KiB); 2-wave ones 12 KiB and 1-wave ones 10.5 KiB, so that eight 1-wave workgroups
    // (19.9 KiB of LDS each, with 16-bit queue entries and records) fit a CU, as many as the
    // registers allow.  A cell whose rows run out moves them to the pool (evict_rows).
    constexpr int NLT = TW >= 4 ? LY::NL : TW == 2 ? LY::NL / 2 : (LY::NL * 7) / 16;
    static_assert(QCAP >= LY::LB, "a slot's non-zero components must fit one round's queue");
    static_assert(LY::LB <= 256 && KB <= 256, "queue entries pack the slot and the local bin in 16 bits");
    static_assert(LY::NL * LY::LB <= 65536, "record addresses are 16-bit LDS payload offsets");
    __shared__ __align__(16) float s_rows[NLT * LY::LB];
    __shared__ FastLds<LY> s_fast[TW];
    __shared__ uint16_t s_q[QCAP];  // (slot << 8 | lidx) of the queued non-zero components
    __shared__ float s_qw[QCAP];
    __shared__ uint32_t s_bid[KB], s_ref[KB], s_cnt[KB], s_off[KB];
    __shared__ uint64_t s_mask[TW];
    __shared__ double s_bsum[TW][NKACC];  // accumulator sums of each wave's batch
    __shared__ uint32_t s_round[2];
    __shared__ uint32_t s_surr;  // learned-physics substeps of the cell
    __shared__ uint32_t s_sdev;  // bits of the largest learned-physics deviation of the cell (>= 0)
    __shared__ uint32_t s_own[NOWN];
    __shared__ uint32_t s_kc[2 * NKC];  // claimed-key cache of the cell (fast_claims)
    __shared__ uint32_t s_home[5 * KB];  // home pool rows of the cell's slots (flush_tables)
    for (int i = threadIdx.x; i < NOWN; i += blockDim.x) s_own[i] = ~0u;
    // The energy tables (edges, ranges, logs): 4-wave workgroups (one per CU) copy them into LDS
    // (dynamic LDS, transport_dyn_lds); the 1- and 2-wave shapes read them through L1/L2, since
    // 4 KiB more per workgroup would fit one one-wave workgroup less per CU.
    Tabs T = Tg;
    if constexpr (TW == 4) {
        extern __shared__ float s_tab[];
        const int NE = Tg.NE;
        float* ee = s_tab;
        float* R = ee + NE + 1;
        float* lR = R + NE;
        float* lE = lR + NE;
        for (int i = threadIdx.x; i <= NE; i += blockDim.x) ee[i] = Tg.ee[i];
        for (int i = threadIdx.x; i < NE; i += blockDim.x) {
            R[i] = Tg.R[i];
            lR[i] = Tg.lR[i];
            lE[i] = Tg.lE[i];
        }
        __syncthreads();
        T.ee = ee;
        T.R = R;
        T.lR = lR;
        T.lE = lE;
    }
    const uint32_t tag = 2 * ctl[C_EP] + (FINE ? 1u : 0u);
    if (ctl[C_DRIFT]) P.surr = 0;  // K5 flagged the learned physics in an earlier iteration
    const int l = lane_id(), wv = (int)(threadIdx.x >> 6);
    const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    const uint32_t n = *nlist;
    CellGeo G;
    G.Nx = T.Nx;
    G.NT = T.NT;
    G.NE = T.NE;
    G.dx = T.dx;
    G.dz = T.dz;
    G.hdx = T.dx * 0.5f;
    G.hdz = T.dz * 0.5f;
    G.tmin = T.te[0];
    G.dth = (T.te[T.NT] - G.tmin) / T.NT;
    G.R_on = P.R_on;
    // The slot headers of a workgroup's next cell (psi_in and psi_out block ids, payload
    // references and slot counts; lanes < 32 of each wave) are fetched while the current cell is
    // transported.  Only this cell's workgroup writes its psi_out during the launch.
    uint32_t nx_cell = blockIdx.x < n ? list[blockIdx.x] : 0u;
    uint32_t nx_ns = 0, nx_bid = EMPTY, nx_ref = 0, nx_ons = 0, nx_obid = EMPTY, nx_oref = 0;
    if (blockIdx.x < n) {
        nx_ns = in.ns[nx_cell];
        nx_ons = out.ns[nx_cell];
        if (l < KB) {
            nx_bid = in.bid[(size_t)nx_cell * KB + l];
            nx_ref = in.sref[(size_t)nx_cell * KB + l];
            nx_obid = out.bid[(size_t)nx_cell * KB + l];
            nx_oref = out.sref[(size_t)nx_cell * KB + l];
        }
    }
    SM2D_STAMP_START();
    for (uint32_t li = blockIdx.x; li < n; li += gridDim.x) {
        SM2D_STAMP_RESTART();
        const int cell = (int)nx_cell;
        const uint32_t ns = nx_ns, c_bid = nx_bid, c_ref = nx_ref, c_ons = nx_ons, c_obid = nx_obid,
                       c_oref = nx_oref;
        if (li + gridDim.x < n) {
            nx_cell = list[li + gridDim.x];
            nx_ns = in.ns[nx_cell];
            nx_ons = out.ns[nx_cell];
            if (l < KB) {
                nx_bid = in.bid[(size_t)nx_cell * KB + l];
                nx_ref = in.sref[(size_t)nx_cell * KB + l];
                nx_obid = out.bid[(size_t)nx_cell * KB + l];
                nx_oref = out.sref[(size_t)nx_cell * KB + l];
            }
        }
        set_cell_geo(G, T, cell, ltag, tag);
        // the cell's seven cumulative tallies, one lane each (four FP64, three FP32 arrays), are
        // read now and written back when the cell ends (no other workgroup writes them)
        double tal_d = 0.0;
        float tal_f = 0.0f;
        if (wv == 0) {
            if (l < 4) tal_d = (l == 0 ? tal.edep : l == 1 ? tal.E_cut : l == 2 ? tal.E_nuc : tal.E_bnd)[cell];
            else if (l < 7) tal_f = (l == 4 ? tal.w_cut : l == 5 ? tal.w_nuc : tal.w_bnd)[cell];
        }
        if (threadIdx.x == 0) {
            s_surr = 0;
            s_sdev = 0;
        }
        if (threadIdx.x < NKC) s_kc[threadIdx.x] = EMPTY;
        if (threadIdx.x < ns) {
            s_bid[threadIdx.x] = c_bid;
            s_ref[threadIdx.x] = c_ref;
        }
        SM2D_STAMP(0);
        // wave 0 owns the tables and the accumulators of the cell
        WTabs<LY> W;
        uint32_t comps = 0;  // components of the cell (wave 0)
        double accd = 0.0;
        float accf = 0.0f;
        if (wv == 0) {
            W.n0 = c_ons;
            W.own0 = W.n0;
            W.b0 = (l < KB && (uint32_t)l < W.n0) ? c_obid : EMPTY;
            W.r0 = (l < KB && (uint32_t)l < W.n0) ? c_oref : 0u;
            W.b1 = W.b2 = W.b3 = W.b4 = EMPTY;
            W.r1 = W.r2 = W.r3 = W.r4 = 0u;
            W.n1 = W.n2 = W.n3 = W.n4 = 0u;
            W.used = 0;
            W.cap = (uint32_t)NLT;
            W.rows = s_rows;
            W.orig = 0u;
            W.stage_own(out);  // e.g. the coarse gather into a fine cell
            for (int i = l; i < 5 * KB; i += 64) s_home[i] = NOHOME;
            if ((uint32_t)l < W.used) s_home[l] = W.orig;  // psi_out slots staged from their payloads
        }
        __syncthreads();
        SM2D_STAMP(1);
        uint32_t s0 = 0;
        while (s0 < ns) {
            // non-zero components per slot (the payloads of SU slots per wave are fetched at once);
            // when the remaining slots fit SR per wave, their payloads are fetched once, kept in
            // registers and queued from there
            constexpr int SU = LY::PL >= 8 ? 1 : 8 / LY::PL;
            constexpr int SR = LY::PL >= 32 ? 1 : 32 / LY::PL;
            const bool one = ns - s0 <= (uint32_t)(SR * TW);
            float vr[SR][LY::PL];
            if (one) {
#pragma unroll
                for (int u = 0; u < SR; ++u) {
                    const uint32_t su = s0 + (uint32_t)wv + (uint32_t)(u * TW);
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) vr[u][k] = 0.0f;
                    if (su < ns && s_bid[su] != EMPTY) {
                        const float* vin = in.pool + (size_t)s_ref[su] * LY::LB;
#pragma unroll
                        for (int k = 0; k < LY::PL; ++k) vr[u][k] = vin[l + 64 * k];
                    }
                }
#pragma unroll
                for (int u = 0; u < SR; ++u) {
                    const uint32_t su = s0 + (uint32_t)wv + (uint32_t)(u * TW);
                    uint32_t cnt = 0;
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) cnt += (uint32_t)__popcll(__ballot(vr[u][k] > 0.0f));
                    if (l == 0 && su < ns) s_cnt[su] = cnt;
                }
            }
            for (uint32_t s = s0 + wv; !one && s < ns; s += SU * TW) {
                float v[SU][LY::PL];
#pragma unroll
                for (int u = 0; u < SU; ++u) {
                    const uint32_t su = s + (uint32_t)(u * TW);
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) v[u][k] = 0.0f;
                    if (su < ns && s_bid[su] != EMPTY) {
                        const float* vin = in.pool + (size_t)s_ref[su] * LY::LB;
#pragma unroll
                        for (int k = 0; k < LY::PL; ++k) v[u][k] = vin[l + 64 * k];
                    }
                }
#pragma unroll
                for (int u = 0; u < SU; ++u) {
                    const uint32_t su = s + (uint32_t)(u * TW);
                    uint32_t cnt = 0;
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) cnt += (uint32_t)__popcll(__ballot(v[u][k] > 0.0f));
                    if (l == 0 && su < ns) s_cnt[su] = cnt;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) {  // the slots of this round: as many as fit the queue
                uint32_t off = 0, s1 = s0;
                while (s1 < ns && off + s_cnt[s1] <= (uint32_t)QCAP) {
                    s_off[s1] = off;
                    off += s_cnt[s1];
                    ++s1;
                }
                s_round[0] = s1;
                s_round[1] = off;
            }
            __syncthreads();
            const uint32_t s1 = s_round[0], total = s_round[1];
            if (one) {
#pragma unroll
                for (int u = 0; u < SR; ++u) {
                    const uint32_t su = s0 + (uint32_t)wv + (uint32_t)(u * TW);
                    if (su < s1) {
                        uint32_t pos = s_off[su];
#pragma unroll
                        for (int k = 0; k < LY::PL; ++k) {
                            const float w = vr[u][k];
                            const uint64_t mk = __ballot(w > 0.0f);
                            if (w > 0.0f) {
                                const uint32_t p = pos + (uint32_t)__popcll(mk & below);
                                s_q[p] = (uint16_t)((su << 8) | (uint32_t)(l + 64 * k));
                                s_qw[p] = w;
                            }
                            pos += (uint32_t)__popcll(mk);
                        }
                    }
                }
            }
            for (uint32_t s = s0 + wv; !one && s < s1; s += SU * TW) {
                float v[SU][LY::PL];
#pragma unroll
                for (int u = 0; u < SU; ++u) {
                    const uint32_t su = s + (uint32_t)(u * TW);
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) v[u][k] = 0.0f;
                    if (su < s1 && s_bid[su] != EMPTY) {
                        const float* vin = in.pool + (size_t)s_ref[su] * LY::LB;
#pragma unroll
                        for (int k = 0; k < LY::PL; ++k) v[u][k] = vin[l + 64 * k];
                    }
                }
#pragma unroll
                for (int u = 0; u < SU; ++u) {
                    const uint32_t su = s + (uint32_t)(u * TW);
                    if (su >= s1) break;
                    uint32_t pos = s_off[su];
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) {
                        const float w = v[u][k];
                        const uint64_t mk = __ballot(w > 0.0f);
                        if (w > 0.0f) {
                            const uint32_t p = pos + (uint32_t)__popcll(mk & below);
                            s_q[p] = (uint16_t)((su << 8) | (uint32_t)(l + 64 * k));
                            s_qw[p] = w;
                        }
                        pos += (uint32_t)__popcll(mk);
                    }
                }
            }
            __syncthreads();
            SM2D_STAMP(2);
            const uint32_t nb = (total + 63) / 64;
            for (uint32_t b0 = 0; b0 < nb; b0 += TW) {
                const uint32_t b = b0 + (uint32_t)wv;
                if (b < nb) {  // physics of batch b by wave wv
                    const uint32_t i = b * 64 + (uint32_t)l;
                    const bool live = i < total;
                    const uint32_t it = live ? s_q[i] : 0u;
                    const uint32_t slot = it >> 8, bid = s_bid[slot];
                    const Learned Lx = learned_for<LY, FINE, INFER>(T, P, G, live, live ? s_qw[i] : 0.0f,
                                                             (int)(it & 0xFFu), bid & 0xFFF, (bid >> 12) & 0xFFF);
                    const Comp r = transport_component<LY, FINE>(T, P, G, live, live ? s_qw[i] : 0.0f,
                                                                 (int)(it & 0xFFu), slot, bid & 0xFFF,
                                                                 (bid >> 12) & 0xFFF, Lx SM2D_SUBA(sc_));
                    const LaneOut o = fast_prepare<LY, FINE>(s_fast[wv], r, live, G.ix, G.iz, G.Nx, T.Nz);
                    s_fast[wv].nkey[l] = o.nkey;
                    s_fast[wv].nrec[l] = o.nrec;
                    SM2D_SUB(5);
                    const double bsum = batch_sum_lane(o.c);
                    if (l < NKACC) s_bsum[wv][l] = bsum;
                    SM2D_SUB(6);
                    const uint64_t m = __ballot(live);
                    const uint32_t nsur = (uint32_t)__popcll(__ballot(live && r.kind == 2 && r.surr));
                    if (l == 0) s_mask[wv] = m;
                    if (FINE && l == 0 && nsur) atomicAdd(&s_surr, nsur);
                    if (FINE && G.monitor) {  // non-negative floats order as their bits
                        uint32_t dv = live ? __float_as_uint(r.sdev) : 0u;
                        for (int o2 = 32; o2 > 0; o2 >>= 1) dv = max(dv, (uint32_t)__shfl_xor((int)dv, o2));
                        if (l == 0 && dv) atomicMax(&s_sdev, dv);
                    }
                    SM2D_STAMP(3);
                    // lateral emissions into other listed cells' +z buckets: mark the targets for
                    // apply_foreign at the end of the launch
                    const uint32_t fm = (live && r.kind == 2 && r.E_new > 0.1f && r.face == 0 && G.iz + 1 < T.Nz)
                                            ? r.fmask : 0u;
                    if (__ballot(fm != 0u)) {
                        for (uint32_t b = fm; b; b &= b - 1) {
                            const int q = __ffs(b) - 1;
                            const uint32_t pT = lpos[cell + q - 5];
                            if (atomicExch(&bk.ftag[pT], tag) != tag) bk.flist[atomicAdd(&ctl[C_FNT], 1u)] = pT;
                        }
                        const uint32_t nc = (uint32_t)__popc(fm);
                        uint32_t tot = nc;
                        for (int o2 = 32; o2 > 0; o2 >>= 1) tot += __shfl_xor(tot, o2);
                        if (l == 0) atomicAdd(&flags[F_CROSS], tot);
                    }
                }
                __syncthreads();
                SM2D_STAMP(4);
                if (wv == 0) {  // ordered apply of the batches, in order
                    for (uint32_t bb = b0; bb < nb && bb < b0 + TW; ++bb) {
                        FastLds<LY>& F = s_fast[bb - b0];
                        const uint64_t m = s_mask[bb - b0];
                        const bool fast = fast_claims<LY>(F, W, s_kc, s_home, out, bk, flags) == FC_OK;
                        SM2D_STAMP(5);
                        if (fast) {
                            fast_rounds<LY>(F, W, m, s_own);
                            SM2D_STAMP(6);
                            add_batch_sum(l < NKACC ? s_bsum[bb - b0][l] : 0.0, accd, accf);
                            SM2D_STAMP(11);
                            comps += (uint32_t)__popcll(m);
                        } else {
                            // the batch again, in wave 0, for the serial form
                            const uint32_t i = bb * 64 + (uint32_t)l;
                            const bool live = i < total;
                            const uint32_t it = live ? s_q[i] : 0u;
                            const uint32_t slot = it >> 8, bid = s_bid[slot];
                            const Learned Lx = learned_for<LY, FINE, INFER>(T, P, G, live, live ? s_qw[i] : 0.0f,
                                                                     (int)(it & 0xFFu), bid & 0xFFF,
                                                                     (bid >> 12) & 0xFFF);
                            const Comp r = transport_component<LY, FINE>(T, P, G, live, live ? s_qw[i] : 0.0f,
                                                                         (int)(it & 0xFFu), slot, bid & 0xFFF,
                                                                         (bid >> 12) & 0xFFF, Lx SM2D_SUBA(nullptr));
                            CellAccum mine = {};
                            serial_apply<LY, FINE>(r, m, W, mine, comps, T, out, bk, flags, G);
                            add_batch_sum(batch_sum_lane(mine), accd, accf);
                            SM2D_STAMP(7);
                        }
                    }
                }
                __syncthreads();
            }
            s0 = s1;
        }
        // ---- end of cell: tables and payloads, tallies, per-cell accumulators ----
        if (wv == 0) {
            CellAccum a = {};
            lanes_to_accum(a, accd, accf);
            a.comps = comps;
            SM2D_STAMP(8);
            flush_tables<LY>(W, out, bk, cell, li, s_home, flags);
            SM2D_STAMP(9);
            if (l < 4) {
                double* t = l == 0 ? tal.edep : l == 1 ? tal.E_cut : l == 2 ? tal.E_nuc : tal.E_bnd;
                const double v = l == 0 ? a.edep : l == 1 ? a.E_cut : l == 2 ? a.E_nuc : a.E_bnd;
                t[cell] = tal_d + v;
            } else if (l < 7) {
                float* t = l == 4 ? tal.w_cut : l == 5 ? tal.w_nuc : tal.w_bnd;
                const float v = l == 4 ? a.w_cut : l == 5 ? a.w_nuc : a.w_bnd;
                t[cell] = tal_f + v;
            }
            if (l == 0) {
                double* o = acc + (size_t)li * NACC;
                o[A_EDEP] = a.edep;
                o[A_WCUT] = a.w_cut;
                o[A_ECUT] = a.E_cut;
                o[A_WNUC] = a.w_nuc;
                o[A_ENUC] = a.E_nuc;
                o[A_WBND] = a.w_bnd;
                o[A_EBND] = a.E_bnd;
                o[A_DROPW] = a.drop_w;
                o[A_DROPE] = a.drop_E;
                o[A_CLWIN] = a.cl_win;
                o[A_CLWOUT] = a.cl_wout;
                o[A_CLEIN] = a.cl_Ein;
                o[A_CLEOUT] = a.cl_Eout;
                o[A_COMPS] = (double)a.comps;
                o[A_UPD] = (double)((size_t)ns * LY::LB);
                o[A_SURR] = (double)s_surr;
                o[A_SDEV] = (double)__uint_as_float(s_sdev);
            }
            SM2D_STAMP(10);
            SM2D_STAMP_CELL();
        }
        __syncthreads();
    }
    SM2D_STAMP_FLUSH();
}

// The phase's lateral emissions into other listed cells' +z buckets (apply_foreign), after the
// transport launch that marked them: one wavefront; returns at once when nothing was marked.
template <class LY, bool FINE, bool INFER>
__global__ __launch_bounds__(64) void k_foreign(Tabs T, Tp P, Psi in, Bkt bk, const uint32_t* __restrict__ list,
                                                const uint32_t* __restrict__ ltag, uint32_t* ctl, uint32_t* flags) {
    const uint32_t nf = ctl[C_FNT];
    if (ctl[C_DONE] || nf == 0) return;
    const uint32_t tag = 2 * ctl[C_EP] + (FINE ? 1u : 0u);
    if (ctl[C_DRIFT]) P.surr = 0;
    CellGeo G;
    G.Nx = T.Nx;
    G.NT = T.NT;
    G.NE = T.NE;
    G.dx = T.dx;
    G.dz = T.dz;
    G.hdx = T.dx * 0.5f;
    G.hdz = T.dz * 0.5f;
    G.tmin = T.te[0];
    G.dth = (T.te[T.NT] - G.tmin) / T.NT;
    G.R_on = P.R_on;
    apply_foreign<LY, FINE, INFER>(T, P, G, in, bk, list, ltag, tag, nf, flags);
    if (threadIdx.x == 0) ctl[C_FNT] = 0u;
}

// Destination set of a phase: the neighbours of the listed cells whose bucket toward them is not
// empty (the gather would add nothing to the others, and a cell that receives nothing and is not
// listed keeps an empty psi_out), deduplicated with a
// per-cell tag and listed in NSH segments of seg_cap entries, each with a counter on a line of
// its own (seg[s * SH_STRIDE]); destination d goes to the segment of a multiplicative hash of d
// (columns and rows of cells spread over all segments), so the segments fill evenly at every
// size and the reservations of a wavefront hit different lines.  The occupancy
// bit of psi_out is set here.  The order of the lists affects no result.
static_assert(NSH == 64, "k_dest hashes cells to 64 segments");
__global__ __launch_bounds__(256) void k_dest(const uint32_t* __restrict__ list, const uint32_t* nlist, int Nx,
                                              int Nz, int Ns, const uint32_t* __restrict__ bkn, uint32_t* dtag,
                                              const uint32_t* ctl, int fine, uint32_t* dlist, uint32_t* seg,
                                              uint32_t seg_cap, uint32_t* bitmap, uint32_t* flags) {
    if (ctl[C_DONE]) return;
    const uint32_t n = *nlist;
    const uint32_t tag = 2 * ctl[C_EP] + (uint32_t)fine;
    const uint32_t nthr = gridDim.x * blockDim.x;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < 4u * n; e += nthr) {
        int d = -1;
        {
            const int c = (int)list[e >> 2], lc = c % Ns;
            const int ix = lc % Nx, iz = lc / Nx;
            switch (e & 3u) {
                case 0: d = iz + 1 < Nz ? c + Nx : -1; break;
                case 1: d = iz > 0 ? c - Nx : -1; break;
                case 2: d = ix + 1 < Nx ? c + 1 : -1; break;
                default: d = ix > 0 ? c - 1 : -1; break;
            }
        }
        if (d < 0 || bkn[e] == 0u) continue;  // bucket e & 3 of listed cell e >> 2 lands in d
        const uint32_t old = dtag[d];
        if (old == tag || atomicCAS(&dtag[d], old, tag) != old) continue;
        const uint32_t s = ((uint32_t)d * 2654435761u) >> 26;  // multiplicative hash, NSH = 64
        const uint32_t p = atomicAdd(seg + s * SH_STRIDE, 1u);
        if (p < seg_cap) dlist[(size_t)s * seg_cap + p] = (uint32_t)d;
        else atomicOr(&flags[F_DEST_FULL], 1u);
        atomicOr(&bitmap[(uint32_t)d >> 5], 1u << ((uint32_t)d & 31));
    }
}

// K4: gather the neighbours' buckets into psi_out of one destination (faces 0..3, bucket slots
// in order), one wavefront per destination.  Workgroup b reads segment b mod NSH of the
// destination list (gridDim is a multiple of NSH).  The direct form (below) adds each block's
// entries in registers.  The staged form, for a destination whose own table would overflow or
// whose new rows the pool cannot give: the destination's payloads are staged in LDS rows
// (existing ones loaded, new ones zeroed), the adds happen in LDS in the serial order, and the
// new payloads get their pool slots with one reservation when the destination is done; the next
// bucket payload is fetched while the current one is added.  COUNT (profiling launches only)
// also adds the destination's non-zero bucket values to moved (NSH shard counters).  KIB: LDS of
// payload rows per destination for the staged form (16, 8 or 4 KiB; the tuner's K4 shape).
// Fewer rows let more wavefronts share a CU; a destination that needs more rows adds into
// global payloads, in the same order, with the same results.
template <class LY, bool COUNT, int KIB>
__global__ __launch_bounds__(64) void k_gather(Tabs T, Psi out, Bkt bk, const uint32_t* __restrict__ dlist,
                                               const uint32_t* __restrict__ seg, uint32_t seg_cap,
                                               const uint32_t* __restrict__ ltag,
                                               const uint32_t* __restrict__ lpos, const uint32_t* ctl, int fine,
                                               double* k4drop_w, double* k4drop_E, uint32_t* flags,
                                               uint32_t* moved) {
    if (ctl[C_DONE]) return;
    constexpr int NLG = KIB * 256 / LY::LB;
    __shared__ __align__(16) float s_rows[NLG * LY::LB];
    const uint32_t tag = 2 * ctl[C_EP] + (uint32_t)fine;
    const int l = lane_id();
    const uint64_t below = (l == 0) ? 0ull : (~0ull >> (64 - l));
    const int Nx = T.Nx, Nz = T.Nz;
    const uint32_t sgi = blockIdx.x % NSH, sstride = gridDim.x / NSH;
    const uint32_t n = min(seg[sgi * SH_STRIDE], seg_cap);  // k_dest flags an overflow
    for (uint32_t i = blockIdx.x / NSH; i < n; i += sstride) {
        const int d = (int)dlist[(size_t)sgi * seg_cap + i];
        const int lc = d % T.Ns;
        const int ix = lc % Nx, iz = lc / Nx;
        // the destination's listed neighbours (face f: its bucket f lands here) and their positions
        int srcs[4];
        uint32_t pos[4];
#pragma unroll
        for (int face = 0; face < 4; ++face) {
            int src = -1;
            if (face == 0 && iz > 0) src = d - Nx;
            if (face == 1 && iz + 1 < Nz) src = d + Nx;
            if (face == 2 && ix > 0) src = d - 1;
            if (face == 3 && ix + 1 < Nx) src = d + 1;
            if (src >= 0 && ltag[src] != tag) src = -1;
            srcs[face] = src;
            pos[face] = src >= 0 ? lpos[src] : 0u;
        }
        // the destination's own psi_out table (lane s < 32: slot s)
        const uint32_t n0 = out.ns[d];
        const uint32_t ob = (l < 32 && (uint32_t)l < n0) ? out.bid[(size_t)d * KB + l] : EMPTY;
        const uint32_t orf = (l < 32 && (uint32_t)l < n0) ? out.sref[(size_t)d * KB + l] : 0u;
        // the four faces' bucket tables (lane s < 32: slot s)
        uint32_t fb[4], fr[4], nf[4];
#pragma unroll
        for (int face = 0; face < 4; ++face) {
            nf[face] = 0;
            fb[face] = EMPTY;
            fr[face] = 0;
            if (srcs[face] >= 0) {
                const uint32_t base = (pos[face] * 4 + face) * KB;  // source face == receiving face index
                nf[face] = bk.n[pos[face] * 4 + face];
                if (l < 32 && (uint32_t)l < nf[face]) {
                    fb[face] = bk.bid[base + l];
                    fr[face] = bk.sref[base + l];
                }
            }
        }
        double dw = 0.0, de = 0.0;
        uint32_t nmoved = 0;  // non-zero bucket values of this destination (profiling only)
        // Direct form: the entries grouped by block, groups in order of first appearance.  A block
        // occurs at most once per bucket table, so a group holds at most four entries (one per
        // face, faces in order); the group's payload (existing or new) and its entries are read
        // together and added in registers in the serial order, then written once.  Taken when
        // the own table has room for every new block (no closest-block fallback) and the pool
        // has the new rows; otherwise the staged form below.
        uint64_t rem[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) rem[f] = __ballot(l < 32 && (uint32_t)l < nf[f] && fb[f] != EMPTY);
        uint32_t nnew = 0;
        {
            uint64_t r2[4] = {rem[0], rem[1], rem[2], rem[3]};
            while (r2[0] | r2[1] | r2[2] | r2[3]) {
                const int f0 = r2[0] ? 0 : r2[1] ? 1 : r2[2] ? 2 : 3;
                const uint32_t fbf = f0 == 0 ? fb[0] : f0 == 1 ? fb[1] : f0 == 2 ? fb[2] : fb[3];
                const uint64_t rf = f0 == 0 ? r2[0] : f0 == 1 ? r2[1] : f0 == 2 ? r2[2] : r2[3];
                const uint32_t b = rlu(fbf, __ffsll((unsigned long long)rf) - 1);
#pragma unroll
                for (int f = 0; f < 4; ++f) r2[f] &= ~__ballot(fb[f] == b);
                if (!__ballot(l < 32 && (uint32_t)l < n0 && ob == b)) ++nnew;
            }
        }
        bool direct = n0 + nnew <= (uint32_t)KB;
        uint32_t nbase = 0;
        if (direct && nnew) {
            if (l == 0) nbase = pool_take(out.pcount, out.pcap, nnew, blockIdx.x);
            nbase = rlu(nbase, 0);
            direct = nbase != POOL_FULL;  // the staged form flags the full pool and books the drops
        }
        if (direct) {
            uint32_t nb = n0, own_b = ob, own_r = orf;
            while (rem[0] | rem[1] | rem[2] | rem[3]) {
                const int f0 = rem[0] ? 0 : rem[1] ? 1 : rem[2] ? 2 : 3;
                const uint32_t fbf = f0 == 0 ? fb[0] : f0 == 1 ? fb[1] : f0 == 2 ? fb[2] : fb[3];
                const uint64_t rf = f0 == 0 ? rem[0] : f0 == 1 ? rem[1] : f0 == 2 ? rem[2] : rem[3];
                const uint32_t b = rlu(fbf, __ffsll((unsigned long long)rf) - 1);
                const uint64_t fo = __ballot(l < 32 && (uint32_t)l < n0 && ob == b);
                uint32_t ref;
                if (fo) {
                    ref = rlu(orf, __ffsll((unsigned long long)fo) - 1);
                } else {
                    ref = nbase + (nb - n0);
                    own_b = (uint32_t)l == nb ? b : own_b;
                    own_r = (uint32_t)l == nb ? ref : own_r;
                    ++nb;
                }
                float acc[LY::PL], v[4][LY::PL];
#pragma unroll
                for (int k = 0; k < LY::PL; ++k) acc[k] = 0.0f;
                if (fo) {
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) acc[k] = out.pool[(size_t)ref * LY::LB + l + 64 * k];
                }
                uint32_t has = 0;
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    const uint64_t gm = rem[f] & __ballot(fb[f] == b);
                    rem[f] &= ~gm;
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) v[f][k] = 0.0f;
                    if (gm) {
                        has |= 1u << f;
                        const uint32_t br = rlu(fr[f], __ffsll((unsigned long long)gm) - 1);
#pragma unroll
                        for (int k = 0; k < LY::PL; ++k) v[f][k] = bk.pool[(size_t)br * LY::LB + l + 64 * k];
                    }
                }
#pragma unroll
                for (int f = 0; f < 4; ++f) {
                    if (!((has >> f) & 1u)) continue;
#pragma unroll
                    for (int k = 0; k < LY::PL; ++k) {
                        if constexpr (COUNT) nmoved += (uint32_t)__popcll(__ballot(v[f][k] > 0.0f));
                        if (v[f][k] > 0.0f) acc[k] += v[f][k];
                    }
                }
#pragma unroll
                for (int k = 0; k < LY::PL; ++k) out.pool[(size_t)ref * LY::LB + l + 64 * k] = acc[k];
            }
            if (l < 32 && (uint32_t)l >= n0 && (uint32_t)l < nb) {
                out.bid[(size_t)d * KB + l] = own_b;
                out.sref[(size_t)d * KB + l] = own_r;
            }
            if (l == 0 && nb != n0) out.ns[d] = nb;
        } else {
            WTabs<LY> W;
            W.n0 = n0;
            W.own0 = n0;
            W.b0 = ob;
            W.r0 = orf;
            W.b1 = W.b2 = W.b3 = W.b4 = EMPTY;
            W.r1 = W.r2 = W.r3 = W.r4 = 0u;
            W.n1 = W.n2 = W.n3 = W.n4 = 0u;
            W.used = 0;
            W.cap = (uint32_t)NLG;
            W.rows = s_rows;
            W.orig = 0u;
            W.stage_own(out);
            const uint32_t ne = nf[0] + nf[1] + nf[2] + nf[3];
            auto entry = [&](uint32_t e, uint32_t& bid, uint32_t& ref) {
                int f = 0;
                while (e >= nf[f]) {
                    e -= nf[f];
                    ++f;
                }
                const uint32_t bb = f == 0 ? fb[0] : f == 1 ? fb[1] : f == 2 ? fb[2] : fb[3];
                const uint32_t rr = f == 0 ? fr[0] : f == 1 ? fr[1] : f == 2 ? fr[2] : fr[3];
                bid = rlu(bb, (int)e);
                ref = rlu(rr, (int)e);
            };
            float v[LY::PL], vn[LY::PL];
            uint32_t bid = EMPTY, bref = 0;
            if (ne) {
                entry(0, bid, bref);
    #pragma unroll
                for (int k = 0; k < LY::PL; ++k) v[k] = bid != EMPTY ? bk.pool[(size_t)bref * LY::LB + l + 64 * k] : 0.0f;
            }
            for (uint32_t e = 0; e < ne; ++e) {
                uint32_t bidn = EMPTY, brefn = 0;
                if (e + 1 < ne) {
                    entry(e + 1, bidn, brefn);
    #pragma unroll
                    for (int k = 0; k < LY::PL; ++k)
                        vn[k] = bidn != EMPTY ? bk.pool[(size_t)brefn * LY::LB + l + 64 * k] : 0.0f;
                }
                if (bid != EMPTY) {
                    if constexpr (COUNT) {
    #pragma unroll
                        for (int k = 0; k < LY::PL; ++k) nmoved += (uint32_t)__popcll(__ballot(v[k] > 0.0f));
                    }
                    uint32_t sel;
                    const int os = W.claim(0, bid, sel, out.pool, out.pcount, out.pcap, flags, F_PSI_FULL);
                    if (os < 0) {
                        double pw = 0.0, pe = 0.0;
    #pragma unroll
                        for (int k = 0; k < LY::PL; ++k) {
                            if (v[k] > 0.0f) {
                                int lidx = l + 64 * k;
                                int eb = min(max((int)((bid >> 12) & 0xFFF) * LY::EL + LY::le(lidx), 0), T.NE - 1);
                                pw += (double)v[k];
                                pe += (double)ecenter(T.ee, eb) * (double)v[k];
                            }
                        }
                        dw += wave_sum(pw);
                        de += wave_sum(pe);
                    } else {
                        const uint32_t ref = rlu(W.r0, os);
                        if (sel == bid) {
    #pragma unroll
                            for (int k = 0; k < LY::PL; ++k)
                                if (v[k] > 0.0f) W.add(ref, l + 64 * k, v[k], out.pool);
                        } else {
                            // closest-block fallback: remapped bins may coincide, add in lidx order
                            for (int k = 0; k < LY::PL; ++k) {
                                for (int j = 0; j < 64; ++j) {
                                    float w = rlf(v[k], j);
                                    if (w > 0.0f && l == 0) W.add(ref, remap<LY>(j + 64 * k, bid, sel), w, out.pool);
                                }
                            }
                        }
                    }
                }
                bid = bidn;
                bref = brefn;
    #pragma unroll
                for (int k = 0; k < LY::PL; ++k) v[k] = vn[k];
            }
            // staged payloads back to the pool: existing rows in place, new rows with one reservation
            const uint64_t mp = __ballot(l < 32 && (uint32_t)l < W.own0 && (W.r0 & CACHED));
            rows_to_pool<LY>(mp, W.r0, W.orig, W.rows, out.pool);
            const uint64_t mo = __ballot(l < 32 && (uint32_t)l >= W.own0 && (uint32_t)l < W.n0 && (W.r0 & CACHED));
            const uint32_t co = (uint32_t)__popcll(mo);
            uint32_t base = 0;
            if (co) {
                if (l == 0) base = pool_take(out.pcount, out.pcap, co, blockIdx.x);
                base = rlu(base, 0);
            }
            const bool ok = !co || base != POOL_FULL;
            if (!ok && l == 0) atomicOr(&flags[F_PSI_FULL], 1u);
            if (ok) {
                uint32_t ref = W.r0;
                if ((mo >> l) & 1ull) ref = base + (uint32_t)__popcll(mo & below);
                rows_to_pool<LY>(mo, W.r0, ref, W.rows, out.pool);
                if (l < 32 && (uint32_t)l >= W.own0 && (uint32_t)l < W.n0) {
                    out.bid[(size_t)d * KB + l] = W.b0;
                    out.sref[(size_t)d * KB + l] = ref;
                }
                if (l == 0) out.ns[d] = W.n0;
            }
        }
        if (l == 0) {
            k4drop_w[d] += dw;
            k4drop_E[d] += de;
            if (COUNT && nmoved) atomicAdd(moved + (blockIdx.x % NSH) * SH_STRIDE, nmoved);
        }
    }
}

// ------------------------------------------------------------------------------------------
// K5: fixed-order FP64 reductions.
// Stage 1 (grid): per block sums over cells of W_in, E_in (from K1), W_out, E_out (psi_out)
// and the K4 drops (consumed and zeroed).  Stage 2 (one block): block partials in order, then
// the per-listed-cell accumulators of both phases in list order.
// ------------------------------------------------------------------------------------------
enum { R_WIN, R_EIN, R_WOUT, R_EOUT, R_K4W, R_K4E, NR1 };
// Audit vector: [0,NR1) cell sums, [NR1,NR1+NACC) phase sums, then max cell W closure, max
// fine-cell E closure, max learned-physics deviation of the monitored fine cells, the number of
// listed cells of the iteration and the cell-bin updates of its fine phase.
constexpr int NAUD = NR1 + NACC + 5;
constexpr int IDX_SDEV = NR1 + NACC + 2;

// Stage 1a: one wavefront per list entry i: W_in / E_in from K1's per-entry sums of psi_in,
// W_out / E_out of psi_out's i-th occupied cell (coalesced payload reads, FP64 butterfly) and
// the K4 drops booked on that cell (consumed).  Per-entry values go to pe[i * NR1 + k].
// `force` runs the pass after the transport ended (totals of a buffer for the host).
template <class LY>
__global__ __launch_bounds__(64) void k_audit_cells(Tabs T, Psi out, const uint32_t* nin, double* win,
                                                    double* ein, const uint32_t* __restrict__ occ_out,
                                                    const uint32_t* nout, double* k4w, double* k4e, double* pe,
                                                    const uint32_t* ctl, int force, Tp P, uint32_t* maskv,
                                                    uint8_t* flag) {
    if (ctl[C_DONE] && !force) return;
    const uint32_t n_in = nin ? *nin : 0u, n_out = *nout;
    const uint32_t n = max(n_in, n_out);
    const uint32_t ep1 = ctl[C_EP] + 1u;  // the next iteration's epoch
    const int l = lane_id();
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        double r[NR1] = {0, 0, 0, 0, 0, 0};
        if (i < n_in) {
            r[R_WIN] = win[i];
            r[R_EIN] = ein[i];
        }
        if (i < n_out) {
            const uint32_t c = occ_out[i];
            const uint32_t ns = out.ns[c];
            // everything else the cell's entry reads, fetched with its slot headers
            const uint32_t tag = maskv ? maskv[c] : 0u;
            const double dk4w = k4w ? k4w[c] : 0.0, dk4e = k4w ? k4e[c] : 0.0;
            const uint32_t mb = (l < (int)ns) ? out.bid[(size_t)c * KB + l] : EMPTY;
            const uint32_t mr = (l < (int)ns) ? out.sref[(size_t)c * KB + l] : 0u;
            double W, E;
            bool bad, nan;
            k1_sums<LY>(T, out, ns, mb, mr, W, E, bad, nan);
            // a NaN weight makes the totals NaN (the audit fails), as a plain sum would
            r[R_WOUT] = nan ? __builtin_nan("") : W;
            r[R_EOUT] = nan ? __builtin_nan("") : E;
            if (k4w && l == 0) {
                r[R_K4W] = dk4w;
                r[R_K4E] = dk4e;
                if (dk4w != 0.0 || dk4e != 0.0) {
                    k4w[c] = 0.0;
                    k4e[c] = 0.0;
                }
            }
            // psi_out is the next iteration's psi_in with this same occupied list: its K1 decision,
            // list flag and input sums are those the next iteration's K1 would compute (entry i is
            // read above before it is replaced here, by the same wavefront)
            if (maskv) {
                const uint8_t f = k1_decide<LY>(P, out, c, ns, mb, W, bad, ep1, maskv, tag);
                if (l == 0) {
                    flag[i] = f;
                    win[i] = W;
                    ein[i] = E;
                }
            }
        }
        if (l == 0)
            for (int k = 0; k < NR1; ++k) pe[(size_t)i * NR1 + k] = r[k];
    }
}

// Stage 1b: per tile of 256 entries, a fixed tree over the entries' values (workgroup b of g).
__device__ inline void audit_tiles_body(double (*sh)[256], const double* __restrict__ pe, const uint32_t* nin,
                                        const uint32_t* nout, double* part, uint32_t b, uint32_t g) {
    const uint32_t n_in = nin ? *nin : 0u, n_out = *nout;
    const uint32_t n = max(n_in, n_out);
    const uint32_t ntile = (n + 255) / 256;
    for (uint32_t tile = b; tile < ntile; tile += g) {
        const uint32_t i = tile * 256 + threadIdx.x;
        for (int k = 0; k < NR1; ++k) sh[k][threadIdx.x] = i < n ? pe[(size_t)i * NR1 + k] : 0.0;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o)
                for (int k = 0; k < NR1; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x < NR1) part[tile * NR1 + threadIdx.x] = sh[threadIdx.x][0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_audit_tiles(const double* __restrict__ pe, const uint32_t* nin,
                                                     const uint32_t* nout, double* part, const uint32_t* ctl,
                                                     int force) {
    if (ctl[C_DONE] && !force) return;
    __shared__ double sh[NR1][256];
    audit_tiles_body(sh, pe, nin, nout, part, blockIdx.x, gridDim.x);
}

// Stage 1c (long lists): per tile of 256 listed cells of a phase, a fixed tree over the cells'
// accumulators: their NACC sums, then the largest W closure, E closure and learned-width
// deviation, and the cell-bin updates (in the fine phase; an integer, exact in any order).
constexpr int NACCP = NACC + 4;

__device__ inline void acc_tiles_body(double (*sh)[256], const double* __restrict__ acc_c, const uint32_t* nc,
                                      const double* __restrict__ acc_f, const uint32_t* nf, double* ap,
                                      uint32_t tcap, uint32_t b, uint32_t g) {
    for (int ph = 0; ph < 2; ++ph) {
        const double* acc = ph ? acc_f : acc_c;
        const uint32_t n = ph ? *nf : *nc;
        const uint32_t ntile = (n + 255) / 256;
        for (uint32_t tile = b; tile < ntile; tile += g) {
            const uint32_t i = tile * 256 + threadIdx.x;
            double v[NACCP];
            for (int k = 0; k < NACCP; ++k) v[k] = 0.0;
            if (i < n) {
                const double* o = acc + (size_t)i * NACC;
                for (int k = 0; k < NACC; ++k) v[k] = o[k];
                if (o[A_CLWIN] > 0) v[NACC] = fabs(o[A_CLWIN] - o[A_CLWOUT]) / o[A_CLWIN];
                if (ph && o[A_CLEIN] > 0) v[NACC + 1] = fabs(o[A_CLEIN] - o[A_CLEOUT]) / o[A_CLEIN];
                if (ph) {
                    v[NACC + 2] = o[A_SDEV];
                    v[NACC + 3] = o[A_UPD];
                }
            }
            for (int k = 0; k < NACCP; ++k) sh[k][threadIdx.x] = v[k];
            __syncthreads();
            for (int o = 128; o > 0; o >>= 1) {
                if ((int)threadIdx.x < o) {
                    for (int k = 0; k < NACC; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
                    for (int k = NACC; k < NACC + 3; ++k)
                        sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + o]);
                    sh[NACC + 3][threadIdx.x] += sh[NACC + 3][threadIdx.x + o];
                }
                __syncthreads();
            }
            if (threadIdx.x < NACCP) ap[((size_t)ph * tcap + tile) * NACCP + threadIdx.x] = sh[threadIdx.x][0];
            __syncthreads();
        }
    }
}

// The audit tiles (workgroups [0, ga)) and the accumulator tiles (the rest) of an iteration with
// long lists in one launch: the two reductions read different arrays and write different partials.
__global__ __launch_bounds__(256) void k_tiles(const double* __restrict__ pe, const uint32_t* nin,
                                               const uint32_t* nout, double* part, uint32_t ga,
                                               const double* __restrict__ acc_c, const uint32_t* nc,
                                               const double* __restrict__ acc_f, const uint32_t* nf, double* ap,
                                               uint32_t tcap, const uint32_t* ctl) {
    if (ctl[C_DONE]) return;
    __shared__ double sh[NACCP][256];
    if (blockIdx.x < ga) audit_tiles_body(sh, pe, nin, nout, part, blockIdx.x, ga);
    else acc_tiles_body(sh, acc_c, nc, acc_f, nf, ap, tcap, blockIdx.x - ga, gridDim.x - ga);
}

// Stage 2 (one workgroup): the tiles in a fixed order, then the per-listed-cell accumulators of
// both phases in list order (or, for long lists, their tile partials from k_acc_tiles, ap).  Output v: [0,NR1) cell sums, [NR1,NR1+NACC) phase sums,
// [NR1+NACC] max cell W closure, [NR1+NACC+1] max fine-cell E closure.  With hist != null the
// vector goes to row (iteration-1) % hist_rows of hist, else to outv.
__global__ __launch_bounds__(256) void k_audit_final(const double* part, const uint32_t* nin, const uint32_t* nout,
                                                     const double* acc_c, const uint32_t* nc, const double* acc_f,
                                                     const uint32_t* nf, double* outv, double* hist,
                                                     uint32_t hist_rows, uint32_t* ctl, int force,
                                                     const double* __restrict__ pe, Psi nxt,
                                                     const uint32_t* __restrict__ nxt_occ, const uint32_t* nxt_n,
                                                     ListOut L, int drift_check, const double* __restrict__ ap,
                                                     uint32_t tcap) {
    if (ctl[C_DONE] && !force) return;
    if (blockIdx.x > 0) {  // clear the next iteration's output buffer (this iteration's input)
        clear_cells(nxt, nxt_occ, *nxt_n, (blockIdx.x - 1) * blockDim.x + threadIdx.x, (gridDim.x - 1) * blockDim.x);
        return;
    }
    constexpr int NV = NAUD - 2;  // reduced entries; the list size and fine updates are appended
    __shared__ double sh[NV][256];
    double r[NV];
    for (int k = 0; k < NV; ++k) r[k] = 0.0;
    const uint32_t n_in = nin ? *nin : 0u, n_out = *nout;
    const uint32_t ntile = (max(n_in, n_out) + 255) / 256;
    if (pe) {  // short lists: the per-cell sums directly (entry i to thread i mod 256, in order)
        const uint32_t n = max(n_in, n_out);
        for (uint32_t i = threadIdx.x; i < n; i += 256)
            for (int k = 0; k < NR1; ++k) r[k] += pe[(size_t)i * NR1 + k];
    } else {
        for (uint32_t b = threadIdx.x; b < ntile; b += 256)
            for (int k = 0; k < NR1; ++k) r[k] += part[b * NR1 + k];
    }
    double fupd = 0.0;  // fine-phase cell-bin updates: integers, exact in FP64 in any order
    for (int ph = 0; ph < 2; ++ph) {
        const double* acc = ph ? acc_f : acc_c;
        const uint32_t n = acc ? (ph ? *nf : *nc) : 0u;
        if (ap && acc) {
            const uint32_t ntile = (n + 255) / 256;
            for (uint32_t b = threadIdx.x; b < ntile; b += 256) {
                const double* o = ap + ((size_t)ph * tcap + b) * NACCP;
                for (int k = 0; k < NACC; ++k) r[NR1 + k] += o[k];
                r[NR1 + NACC] = fmax(r[NR1 + NACC], o[NACC]);
                r[NR1 + NACC + 1] = fmax(r[NR1 + NACC + 1], o[NACC + 1]);
                r[IDX_SDEV] = fmax(r[IDX_SDEV], o[NACC + 2]);
                fupd += o[NACC + 3];
            }
            continue;
        }
        for (uint32_t i = threadIdx.x; i < n; i += 256) {
            const double* o = acc + (size_t)i * NACC;
            for (int k = 0; k < NACC; ++k) r[NR1 + k] += o[k];
            if (o[A_CLWIN] > 0) r[NR1 + NACC] = fmax(r[NR1 + NACC], fabs(o[A_CLWIN] - o[A_CLWOUT]) / o[A_CLWIN]);
            if (ph && o[A_CLEIN] > 0)
                r[NR1 + NACC + 1] = fmax(r[NR1 + NACC + 1], fabs(o[A_CLEIN] - o[A_CLEOUT]) / o[A_CLEIN]);
            if (ph) {
                r[IDX_SDEV] = fmax(r[IDX_SDEV], o[A_SDEV]);
                fupd += o[A_UPD];
            }
        }
    }
    __shared__ double sh_fupd[256];
    for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] = r[k];
    sh_fupd[threadIdx.x] = fupd;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            for (int k = 0; k < NR1 + NACC; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
            for (int k = NR1 + NACC; k < NV; ++k) sh[k][threadIdx.x] = fmax(sh[k][threadIdx.x], sh[k][threadIdx.x + o]);
            sh_fupd[threadIdx.x] += sh_fupd[threadIdx.x + o];
        }
        __syncthreads();
    }
    double* dst = hist ? hist + (size_t)((ctl[C_IT] - 1u) % hist_rows) * NAUD : outv;
    if (threadIdx.x < NV) dst[threadIdx.x] = sh[threadIdx.x][0];
    if (drift_check && threadIdx.x == 0) {
        // K5 flags the learned physics: a learned width far from the plain one in a monitored
        // cell, a weight balance off by 1e-6 or a non-finite energy balance (source terms, which
        // enter both sides of the balances, left out); the next iteration runs the plain physics
        const double W_in = sh[R_WIN][0];
        const double Wr = sh[R_WOUT][0] + sh[NR1 + A_WCUT][0] + sh[NR1 + A_WNUC][0] + sh[NR1 + A_WBND][0] +
                          fmax(0.0, sh[NR1 + A_DROPW][0] + sh[R_K4W][0]);
        const double Er = sh[R_EOUT][0] + sh[NR1 + A_EDEP][0] + sh[NR1 + A_ECUT][0] + sh[NR1 + A_ENUC][0] +
                          sh[NR1 + A_EBND][0] + fmax(0.0, sh[NR1 + A_DROPE][0] + sh[R_K4E][0]);
        const double werr = fabs(W_in - Wr) / fmax(W_in, 1e-20);
        if (!(werr < 1e-6) || !isfinite(sh[R_EIN][0] - Er) || !(sh[IDX_SDEV][0] <= (double)SURR_DRIFT_TOL))
            ctl[C_DRIFT] = 1u;
    }
    if (threadIdx.x == NV) dst[NV] = (double)((acc_c ? *nc : 0u) + (acc_f ? *nf : 0u));
    if (threadIdx.x == 0) dst[NV + 1] = sh_fupd[0];  // fine-phase updates
    if (gridDim.x > 1) {  // open the next iteration once every thread has read the counters
        __syncthreads();
        if (L.occ) {  // the next iteration's lists from the K1 flags of psi_out (k_audit_cells)
            __shared__ uint32_t wc[2][16];
            __shared__ uint32_t base[2];
            compact_lists(L, n_out, ctl[C_EP] + 1u, ctl[C_IT] + 1u, ctl, wc, base);
            __syncthreads();
        }
        if (threadIdx.x == 0) open_iteration(ctl);
    }
}

// ------------------------------------------------------------------------------------------
// Source injection (k1k6_pipeline.cu:300-551, device_psic.cuh:347-450).  Samples are drawn in
// parallel; their phase-space bins are then applied in sample order by one wavefront, so slot
// claims are deterministic (the reference claims them in thread-race order).
// ------------------------------------------------------------------------------------------
struct InjSample {
    int cell;      // -1 outside the grid
    uint32_t bid;
    int lidx;
    float E;
};

__device__ inline float inj_gauss(unsigned& s) {  // gaussian_sample, k1k6_pipeline.cu:433-447
    s = 1664525u * s + 1013904223u;
    float u1 = (s >> 16) / 65536.0f;
    s = 1664525u * s + 1013904223u;
    float u2 = (s >> 16) / 65536.0f;
    float r = sqrtf_(-2.0f * logf_(fmaxf(u1, 1e-10f)));
    float th = (float)(2.0 * 3.14159265358979323846 * (double)u2);
    return r * cosf_(th);
}

template <class LY>
__device__ inline InjSample inj_bin(const Tabs& T, int cell, float theta, float E, float x, float z) {
    InjSample o;
    o.cell = cell;
    o.E = E;
    x = fmaxf(0.0f, fminf(x, T.dx));
    z = fmaxf(0.0f, fminf(z, T.dz));
    int xs = min(max((int)((((x - T.dx * 0.5f) / T.dx) + 0.5f) * LY::XS), 0), LY::XS - 1);
    int zs = min(max((int)((((z - T.dz * 0.5f) / T.dz) + 0.5f) * LY::ZS), 0), LY::ZS - 1);
    const int NT = T.NT, NE = T.NE;
    float tmin = T.te[0], tmax = T.te[NT], Emin = T.ee[0], Emax = T.ee[NE];
    theta = fmaxf(tmin, fminf(theta, tmax));
    E = fmaxf(Emin, fminf(E, Emax));
    float dth = (tmax - tmin) / NT;
    float tc = (theta - tmin) / dth;
    int tb = (int)tc;
    float ft = tc - tb;
    int eb = find_bin(T.ee, NE, E);
    float lo = T.ee[eb], hi = T.ee[eb + 1];
    float fe = (hi - lo) > 1e-10f ? (E - lo) / (hi - lo) : 0.0f;
    if (tb >= NT - 1) { tb = NT - 1; ft = 0.0f; }
    if (eb >= NE - 1) { eb = NE - 1; fe = 0.0f; }
    if (ft >= 0.5f && tb < NT - 1) tb++;
    if (fe >= 0.5f && eb < NE - 1) eb++;
    o.bid = ((uint32_t)(tb / LY::TL) & 0xFFF) | (((uint32_t)(eb / LY::EL) & 0xFFF) << 12);
    o.lidx = LY::enc(tb % LY::TL, eb % LY::EL, xs, zs);
    return o;
}

__device__ inline bool spot_is_gaussian(const sm2d_spot& sp) {  // gpu_transport_wrapper.cu:240
    return (sp.n_samples > 1) || (sp.sigma_x > 0) || (sp.sigma_theta > 0);
}

// Samples of all spots: out[spot * nmax + i]; a Gaussian spot draws n_samples
// (inject_gaussian_source_kernel, k1k6_pipeline.cu:451-551), a pencil spot one entry
// (inject_source_kernel pencil branch, k1k6_pipeline.cu:325-343).
template <class LY>
__global__ void k_inject_samples(Tabs T, const sm2d_spot* __restrict__ spots, int S, int nmax, InjSample* out) {
    const int total = S * nmax;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
        const int s = idx / nmax, i = idx - s * nmax;
        const sm2d_spot sp = spots[s];
        InjSample o;
        o.cell = -2;
        o.bid = EMPTY;
        o.lidx = 0;
        o.E = 0.0f;
        const int base = s * T.Ns;
        if (spot_is_gaussian(sp)) {
            if (i < sp.n_samples) {
                unsigned seed = sp.random_seed + (unsigned)i;
                float xs = sp.x0 + sp.sigma_x * inj_gauss(seed);
                float ts = sp.theta0 + sp.sigma_theta * inj_gauss(seed);
                float Es = sp.E0 + sp.sigma_E * inj_gauss(seed);
                Es = fmaxf(T.ee[0], fminf(Es, T.ee[T.NE]));
                ts = fmaxf(T.te[0], fminf(ts, T.te[T.NT]));
                float xr = xs - T.x_min, zr = sp.z0 - T.z_min;
                int ix = (int)(xr / T.dx), iz = (int)(zr / T.dz);
                if (ix < 0 || ix >= T.Nx || iz < 0 || iz >= T.Nz) {
                    o.cell = -1;
                    o.E = Es;
                } else {
                    o = inj_bin<LY>(T, base + iz * T.Nx + ix, ts, Es, xr - ix * T.dx, zr - iz * T.dz);
                }
            }
        } else if (i == 0) {
            float xr = sp.x0 - T.x_min, zr = sp.z0 - T.z_min;
            int sx = (int)(xr / T.dx), sz = (int)(zr / T.dz);
            sx = sx < 0 ? 0 : (sx >= T.Nx ? T.Nx - 1 : sx);
            sz = sz < 0 ? 0 : (sz >= T.Nz ? T.Nz - 1 : sz);
            o = inj_bin<LY>(T, base + sz * T.Nx + sx, sp.theta0, sp.E0, xr - sx * T.dx, zr - sz * T.dz);
        }
        out[idx] = o;
    }
}

// One wavefront per spot applies its samples in order (deterministic slot claims; the
// reference claims them in thread-race order).  acc[spot*6 + k]: W in / out-of-grid / slot
// drop, E in / out-of-grid / slot drop.  Touched cells are marked in the occupancy bitmap.
template <class LY>
__global__ __launch_bounds__(64) void k_inject_apply(Psi in, const sm2d_spot* __restrict__ spots,
                                                     const InjSample* smp, int nmax, double* acc,
                                                     uint32_t* bitmap, uint32_t* flags) {
    const int l = lane_id();
    const int spi = blockIdx.x;
    const sm2d_spot sp = spots[spi];
    const bool gauss = spot_is_gaussian(sp);
    const int n = gauss ? sp.n_samples : 1;
    const float w = gauss ? sp.W_total / sp.n_samples : sp.W_total;
    const InjSample* s = smp + (size_t)spi * nmax;
    double Wi = 0, Wo = 0, Ws = 0, Ei = 0, Eo = 0, Es = 0;
    for (int i = 0; i < n; ++i) {
        const InjSample q = s[i];
        if (q.cell < 0) {
            Wo += (double)w;
            Eo += (double)(q.E * w);
            continue;
        }
        if (!(w > 0.0f)) {  // device_psic_inject_source rejects non-positive weights
            Ws += (double)w;
            Es += (double)(q.E * w);
            continue;
        }
        uint32_t nn = in.ns[q.cell];
        uint32_t mb = (l < 32 && (uint32_t)l < nn) ? in.bid[(size_t)q.cell * KB + l] : EMPTY;
        uint32_t mr = (l < 32 && (uint32_t)l < nn) ? in.sref[(size_t)q.cell * KB + l] : 0u;
        int slot = -1;
        uint64_t hit = __ballot(l < 32 && mb == q.bid);
        if (hit) {
            slot = __ffsll((unsigned long long)hit) - 1;
        } else if (nn < (uint32_t)KB) {  // device_psic_find_or_allocate_slot: no fallback
            uint32_t sel;
            slot = table_find_or_claim(mb, mr, nn, q.bid, sel, in.pool, in.pcount, in.pcap,
                                       in.bid + (size_t)q.cell * KB, in.sref + (size_t)q.cell * KB, LY::LB,
                                       flags, F_PSI_FULL);
            if (l == 0 && slot >= 0) {
                in.ns[q.cell] = nn;
                atomicOr(&bitmap[(uint32_t)q.cell >> 5], 1u << ((uint32_t)q.cell & 31));
            }
        }
        if (slot >= 0) {
            uint32_t ref = rlu(mr, slot);
            if (l == 0) in.pool[(size_t)ref * LY::LB + q.lidx] += w;
            Wi += (double)w;
            Ei += (double)(q.E * w);
        } else {
            Ws += (double)w;
            Es += (double)(q.E * w);
        }
        __threadfence_block();  // the next sample's slot search reads what lane 0 wrote
    }
    if (l == 0) {
        if (!gauss) {  // the wrapper books the pencil source as E0 * W (gpu_transport_wrapper.cu:328-333)
            Wi = (double)w;
            Ei = (double)sp.E0 * (double)w;
            Wo = Ws = Eo = Es = 0.0;
        }
        double* a = acc + (size_t)spi * 6;
        a[0] = Wi;
        a[1] = Wo;
        a[2] = Ws;
        a[3] = Ei;
        a[4] = Eo;
        a[5] = Es;
    }
}

// ------------------------------------------------------------------------------------------
// Stage helpers for tests: load / dump of psi_in rows
// ------------------------------------------------------------------------------------------
__global__ void k_math_eval(int which, const float* x, float* y, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = x[i];
    switch (which) {
        case 0: y[i] = expf_(v); break;
        case 1: y[i] = logf_(v); break;
        case 2: y[i] = sinf_(v); break;
        case 3: y[i] = cosf_(v); break;
        case 5:
        case 6: {
            float sn, cs;
            sincosf_(v, sn, cs);
            y[i] = which == 5 ? sn : cs;
            break;
        }
        default: y[i] = erff_(v); break;
    }
}

}  // namespace sm2d
