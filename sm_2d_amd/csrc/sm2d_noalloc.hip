// sm2d_noalloc.hip — K1 ActiveMask + K2 CompactActive on caller-provided device buffers, with
// no device allocation: the entry point sm2d_k1k2_mask_compact_noalloc (include/sm2d_hip.h).
//
// Its scratch (control words, per-tile counts of the ordered compaction and one flag byte per
// cell) is the file-scope device array `scratch` below, of a fixed size: SCRATCH_MAX_CELLS cells
// at most.  Host code never passes the array's name to a kernel: its device address is looked up
// once per device (hipGetSymbolAddress, on the entry point's first call with that device
// current) and that pointer is what the kernels receive.  Every call on a device uses the same
// array, so calls that overlap on one device race; each call clears what must start at zero
// with a kernel of its own (no memset node, so the calls can be captured in a graph).
//
// Included into sm2d_runtime.hip (one translation unit with the kernels it reuses:
// k_flag_count and k_scan<2>).
#pragma once

#include <mutex>

namespace sm2d {

constexpr uint32_t SCRATCH_MAX_CELLS = 1u << 22;                 // 4 194 304 cells
constexpr uint32_t SCRATCH_TILES = SCRATCH_MAX_CELLS / 256;      // per-tile counts of the scan
constexpr size_t SCRATCH_WORDS = 64 + 2 * (size_t)SCRATCH_TILES;  // control words, then counts
constexpr size_t SCRATCH_BYTES = 4 * SCRATCH_WORDS + SCRATCH_MAX_CELLS;  // then a flag per cell
// word layout: [0, NCTL) control block, [16] cell count, [32, 32 + NFLAG) error flags,
// [64, 64 + 2 * tiles) tile counts; byte 4 * SCRATCH_WORDS on: the flags of the cells
constexpr uint32_t SW_N = 16, SW_FLAGS = 32, SW_TILES = 64;

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

// Zeroes the control block and error flags and stores the cell count.
__global__ void k_noalloc_init(uint32_t* sw, uint32_t n) {
    const uint32_t i = threadIdx.x;
    if (i < 64) sw[i] = i == SW_N ? n : 0u;
}

// K1 on the reference's dense layout (DevicePsiC, device_psic.cuh:39-55: block_ids[N * 32],
// values[N * 32 * LB]), K1_ActiveMask (k1_activemask.cu:6-48): one wavefront per cell.  The
// weight test W > w_min is that of the reference's serial FP32 sum; as in k1_decide, an FP64
// sum decides unless it lies within 1e-3 of w_min (or a weight is negative or NaN), where lane 0
// forms the serial sum itself.  flag: 1 active, 2 coarse (mask 0 with weight), 0 nothing.
template <int LB>
__global__ __launch_bounds__(64) void k_dense_k1(const uint32_t* __restrict__ bid, const float* __restrict__ val,
                                                 const uint8_t* __restrict__ prev, uint8_t* mask, uint8_t* flag,
                                                 uint32_t N, int b_on, int b_off, float wmin) {
    constexpr int PL = LB / 64;
    const int l = lane_id();
    for (uint32_t c = blockIdx.x; c < N; c += gridDim.x) {
        const uint32_t mb = l < KB ? bid[(size_t)c * KB + l] : EMPTY;
        const uint32_t bE = (mb >> 12) & 0xFFF;
        const bool on = __ballot(mb != EMPTY && bE <= (uint32_t)b_on) != 0;
        const bool off = __ballot(mb != EMPTY && bE <= (uint32_t)b_off) != 0;
        double W64 = 0.0;
        bool neg = false;
        for (int s = 0; s < KB; ++s) {
            if (rlu(mb, s) == EMPTY) continue;
            const float* v = val + ((size_t)c * KB + s) * LB;
#pragma unroll
            for (int k = 0; k < PL; ++k) {
                const float w = v[l + 64 * k];
                if (w > 0.0f) W64 += (double)w;
                else if (w < 0.0f || w != w) neg = true;
            }
        }
        W64 = wave_sum(W64);
        const bool bad = __ballot(neg) != 0;
        const double wm = (double)wmin;
        bool pass;
        if (!bad && W64 > wm * (1.0 + 1e-3)) pass = true;
        else if (!bad && W64 < wm * (1.0 - 1e-3)) pass = false;
        else {
            float W = 0.0f;  // the reference's serial FP32 sum (k1_activemask.cu:29-41)
            if (l == 0)
                for (int s = 0; s < KB; ++s) {
                    if (bid[(size_t)c * KB + s] == EMPTY) continue;
                    const float* v = val + ((size_t)c * KB + s) * LB;
                    for (int q = 0; q < LB; ++q) W += v[q];
                }
            pass = rlf(W, 0) > wmin;
        }
        if (l == 0) {
            const bool pa = prev != nullptr && prev[c] != 0;
            const bool m = (on || (pa && off)) && pass;
            mask[c] = m ? 1 : 0;
            flag[c] = m ? 1 : (W64 > 0.0 ? 2 : 0);
        }
    }
}

// Ordered scatter of the compaction over cells 0..N-1 (ascending cell order: the canonical form
// of compact_active_list / compact_coarse_list, k1k6_pipeline.cu:589-665); block 0 also stores
// the two list lengths.
__global__ __launch_bounds__(256) void k_dense_emit(const uint8_t* __restrict__ flag, const uint32_t* sw,
                                                    const uint32_t* __restrict__ toff, uint32_t* act, uint32_t* crs,
                                                    uint32_t* counts) {
    const uint32_t* ctl = sw;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counts[0] = ctl[C_NACT];
        counts[1] = ctl[C_NCRS];
    }
    if (ctl[C_DONE]) return;
    __shared__ uint32_t wcnt[2][4];
    const uint32_t n = sw[SW_N];
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
            uint32_t pos = toff[tile * 2 + k];
            for (int j = 0; j < w; ++j) pos += wcnt[k][j];
            pos += __popcll((k == 0 ? ba : bc) & below);
            (k == 0 ? act : crs)[pos] = i;
        }
        __syncthreads();
    }
}

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};  // device address of `scratch` per device
}  // namespace

// The device address of `scratch` on the current device (looked up once per device).
inline int scratch_on_current_device(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return err("no-allocation entry point: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return err(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e), SM2D_ERR_HIP);
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SM2D_OK;
}

}  // namespace sm2d

extern "C" int sm2d_k1k2_mask_compact_noalloc(int Nx, int Nz, int local_bins, const uint32_t* d_block_ids,
                                              const float* d_values, const uint8_t* d_mask_prev, uint8_t* d_mask,
                                              int b_E_fine_on, int b_E_fine_off, float weight_active_min,
                                              uint32_t* d_active, uint32_t* d_coarse, uint32_t* d_counts,
                                              void* stream) {
    using namespace sm2d;
    if (Nx <= 0 || Nz <= 0) return err("sm2d_k1k2_mask_compact_noalloc: bad grid");
    const uint64_t N = (uint64_t)Nx * (uint64_t)Nz;
    if (N > SCRATCH_MAX_CELLS)
        return err("sm2d_k1k2_mask_compact_noalloc: " + std::to_string(N) + " cells exceed the scratch limit of " +
                   std::to_string(SCRATCH_MAX_CELLS));
    if (local_bins != 128 && local_bins != 256)
        return err("sm2d_k1k2_mask_compact_noalloc: local_bins must be 128 or 256", SM2D_ERR_UNSUPPORTED);
    if (!d_block_ids || !d_values || !d_mask || !d_active || !d_coarse || !d_counts)
        return err("sm2d_k1k2_mask_compact_noalloc: null buffer");
    void* base = nullptr;
    int rc = scratch_on_current_device(&base);
    if (rc) return rc;
    uint32_t* sw = (uint32_t*)base;
    uint8_t* flag = (uint8_t*)base + 4 * SCRATCH_WORDS;
    uint32_t* tiles = sw + SW_TILES;
    const hipStream_t s = (hipStream_t)stream;
    const uint32_t n = (uint32_t)N;
    k_noalloc_init<<<1, 64, 0, s>>>(sw, n);
    const int gk = blocks_for(n, 1, 1 << 16);
    if (local_bins == 256)
        k_dense_k1<256><<<gk, 64, 0, s>>>(d_block_ids, d_values, d_mask_prev, d_mask, flag, n, b_E_fine_on,
                                          b_E_fine_off, weight_active_min);
    else
        k_dense_k1<128><<<gk, 64, 0, s>>>(d_block_ids, d_values, d_mask_prev, d_mask, flag, n, b_E_fine_on,
                                          b_E_fine_off, weight_active_min);
    const int gt = blocks_for(n, 256, 4096);
    k_flag_count<<<gt, 256, 0, s>>>(flag, sw + SW_N, tiles, sw);
    k_scan<2><<<1, 1024, 0, s>>>(tiles, sw + SW_N, 256, sw, nullptr, n, sw + SW_FLAGS, 0u);
    k_dense_emit<<<gt, 256, 0, s>>>(flag, sw, tiles, d_active, d_coarse, d_counts);
    HIPCHK(hipGetLastError());
    return SM2D_OK;
}
