// sm2d_runtime.hip — context, iteration orchestration and the C ABI of libsm2d_hip.so.
//
// One context = one problem grid times S spot slots on one device, with its own HIP stream and
// all device memory (persistent across spots and batches).  An iteration mirrors
// run_k1k6_pipeline_transport (src/cuda/k1k6_pipeline.cu:1540-2038):
//   reset psi_out -> K1 mask -> K2 compaction -> [coarse K3 -> K4] -> [fine K3 -> K4] -> K5 -> K6
// Every pass walks the occupied-cell lists, never the grid (except one bitmap word per 32
// cells).  Epoch, counts and the end-of-transport test live on the device, so iterations are
// queued in chunks and the host reads back once per chunk instead of the reference's ~60
// synchronous calls per iteration.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include <hip/hip_ext.h>

#include "../../include/sm2d_hip.h"
#include "sm2d_kernels.hip"

namespace sm2d_host {
void set_error(const std::string& m);
const char* error();
}

using namespace sm2d;

namespace {

int err(const std::string& m, int code = SM2D_ERR_ARG) {
    sm2d_host::set_error(m);
    return code;
}

#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess)                                                         \
            return err(std::string(#x) + ": " + hipGetErrorString(e_), SM2D_ERR_HIP); \
    } while (0)

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    size_t bytes_for(size_t count) const { return sizeof(T) * (count ? count : 1); }
    hipError_t alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        return hipMalloc((void**)&p, sizeof(T) * (count ? count : 1));
    }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
};

enum LayoutId { L_REF = 0, L_CFG = 1 };
constexpr int NV = NAUD;  // audit vector
constexpr int IDX_LIST = NAUD - 2, IDX_FUPD = NAUD - 1;
constexpr int HROWS = 64;           // audit rows per chunk
constexpr int CHUNK = 32;           // iterations queued per host read-back

}  // namespace

// Launch-shape tuner (every choice gives the same results).  Four tuned launches: the coarse and
// fine transports (kinds 0, 1: waves per listed cell, 1, 2 or 4) and the coarse and fine K4
// gathers (kinds 2, 3: LDS payload rows per destination, 16, 8 or 4 KiB).  The choice is made per
// kind and per size class of the launch (floor(log2) of the phase's listed cells in the previous
// chunk): the first chunk of a class runs the three shapes on alternating iterations with events
// on each launch; later chunks use the fastest average.
struct Tuner {
    static constexpr int NK = 4;
    static constexpr int NC = 3;
    static constexpr int NB = 24;
    double ms[NK][NB][NC] = {};
    int cnt[NK][NB][NC] = {};
    bool done[2][NB] = {};    // per phase: both of its kinds explored
    int forced = 0;           // 0 = tuned, else the fixed waves per cell
    int forced_kib = 0;       // 0 = tuned, else the fixed gather rows (KiB)
    static int tw(int c) { return c == 0 ? 1 : c == 1 ? 2 : 4; }
    static int kib(int c) { return c == 0 ? 16 : c == 1 ? 8 : 4; }
    // the shape kept on near-ties: four waves per cell, 16 KiB of gather rows
    static int dflt(int kind) { return kind < 2 ? 2 : 0; }
    static int bucket(uint32_t listed) {
        int b = 0;
        while (b + 1 < NB && (2u << b) <= listed) ++b;
        return b;
    }
    // The fastest shape; the default when it is within 5 % of the fastest, so that near-ties do
    // not flip with timing noise from one process to the next.  Transport lists of 2048 cells or
    // more take one wave per cell: eight cells in flight per CU against one for four waves, and
    // a class's first chunk can mislead (a fine list that has just appeared is cheap per cell in
    // the iterations the tuner times, then grows and becomes the costliest phase; measured on the
    // 1029-bin batch, four waves there took 2.7 ms per launch against 1.9 ms for one).
    static constexpr int ONE_WAVE = 11;  // size class of 2048 listed cells
    int best(int kind, int b) const {
        if (kind < 2 && b >= ONE_WAVE) return 0;
        const int d = dflt(kind);
        int bc = d;
        double bt = 1e300;
        for (int c = 0; c < NC; ++c)
            if (cnt[kind][b][c] && ms[kind][b][c] / cnt[kind][b][c] < bt) {
                bt = ms[kind][b][c] / cnt[kind][b][c];
                bc = c;
            }
        if (cnt[kind][b][d] && ms[kind][b][d] / cnt[kind][b][d] <= 1.05 * bt) bc = d;
        return bc;
    }
};

struct sm2d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    sm2d_grid g{};
    LayoutId layout = L_REF;
    int LB = 256, S = 1, Ns = 0, N = 0;
    uint32_t pcap = 0, lcap = 0, dcap = 0, ocap = 0, nwords = 0;
    uint64_t bytes = 0;  // device bytes of the buffers sm2d_ctx_create allocated
    int batch_n = 0;     // spots of the last sm2d_run_batch (their EdepC stays readable)
    uint64_t skip_chunks = 0, pauses = 0;  // graph chunks without a phase, and their pauses
    bool tables = false;
    DBuf<float> te, ee, R, Sp, lE, lR, lS, feA, feB, sx0, rho, x0r;
    std::vector<float> h_ee, h_R, h_lR, h_lE;  // host copies of the energy tables (launch constants)
    DBuf<uint32_t> bid[2], ns[2], sref[2], pcount[2], occ[2];
    DBuf<float> pool[2];
    int in = 0;
    DBuf<uint32_t> bitmap;
    DBuf<uint32_t> bbid, bn, bsref, bftag, bflist;
    DBuf<float> bpool;
    DBuf<uint32_t> maskv, ltag, lpos, dtag, act, crs, dlist, ctl, tcount, tcount2, flags;
    DBuf<uint32_t> gmoved;  // profiling: bucket values moved per (iteration slot, phase) of a chunk,
                            // in NSH shard counters each
    DBuf<uint8_t> flag;
    DBuf<double> edep, E_cut, E_nuc, E_bnd, win, ein, k4w, k4e, acc_c, acc_f, pe, part, k5v, hist, edep_sum;
    DBuf<double> tsum;  // tally totals: per-workgroup partial sums, then the seven totals
    DBuf<double> accpart;  // tile partials of the listed cells' accumulators, per phase
    DBuf<float> w_cut, w_nuc, w_bnd, edep_f;
    DBuf<sm2d_spot> spots;
    DBuf<InjSample> samples;
    DBuf<double> injacc;
    // learned physics of the fine transport (sm2d_surrogate_attach)
    DBuf<float> surr;
    DBuf<int> rowmat;
    int surr_nodes = 0, surr_nmat = 0, surr_on = 0;  // surr_on: 0 off, 1 table, 2 network inference
    const uint16_t* mlp_wb = nullptr;  // the attached network (inference in the fine transport)
    const float* mlp_p = nullptr;
    float mlp_dict[4] = {};
    float surr_tol = 0.25f;   // validation tolerance of an attached network's table
    float surr_stats[4] = {}; // last validation: nodes checked, flagged, max and RMS deviation
    std::vector<double> src;  // per spot: W in/out/slot, E in/out/slot, E representation loss
    uint32_t hint = 64;       // expected listed cells per phase (launch sizing only)
    uint32_t hint_ph[2] = {64, 64};  // listed cells of the last chunk, coarse / fine
    Tuner tune;
    int cand[4] = {2, 2, 0, 0};       // shapes of the iteration being queued (Tuner kinds)
    std::vector<int> cand_log;        // shapes of the chunk's iterations (4 per iteration)
    uint32_t hint_occ = 64;   // expected occupied cells per buffer (launch sizing only)
    // stage timing (eager launches with events on `stream`)
    // Events are recorded for every queued iteration of a chunk (12 per iteration) and read
    // after the chunk's host synchronisation, so profiling does not change the queueing.
    static constexpr int EPI = 12;
    bool prof = false;
    std::vector<hipEvent_t> ev;
    int ev_it = 0;  // iteration slot within the current chunk
    sm2d_timing timing{};
    bool rec = false;  // events recorded this chunk (profiling or tuner exploration)
    // Captured chunks of iterations (hipGraph), keyed by everything their launches depend on.
    struct ChunkGraph {
        std::vector<uint8_t> key;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    std::vector<ChunkGraph> graphs;
    bool graphs_on = true;
    uint64_t graph_clock = 0, graph_captures = 0, graph_replays = 0;
    void graphs_clear() {
        for (ChunkGraph& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
    void mark(int k) {
        if (rec) (void)hipEventRecord(ev[(size_t)ev_it * EPI + k], stream);
    }

    Tabs tabs() const {
        Tabs t;
        t.Nx = g.Nx; t.Nz = g.Nz; t.N = N; t.NT = g.N_theta; t.NE = g.N_E;
        t.Ns = Ns; t.S = S;
        t.dx = g.dx; t.dz = g.dz; t.x_min = g.x_min; t.z_min = g.z_min;
        t.te = te.p; t.ee = ee.p; t.R = R.p; t.lR = lR.p; t.lE = lE.p;
        t.feA = feA.p; t.feB = feB.p; t.sx0 = sx0.p; t.rho = rho.p; t.x0r = x0r.p;
        t.surr = surr.p; t.rowmat = rowmat.p; t.surr_nodes = surr_nodes;
        t.mlp_wb = mlp_wb; t.mlp_p = mlp_p;
        for (int k = 0; k < 4; ++k) t.mlp_dict[k] = mlp_dict[k];
        return t;
    }
    Psi psi(int k) const {
        Psi p;
        p.bid = bid[k].p; p.ns = ns[k].p; p.sref = sref[k].p; p.pool = pool[k].p;
        p.pcount = pcount[k].p; p.pcap = pcap;
        return p;
    }
    // Bucket tables of a phase: both phases share the pool (the coarse gather has consumed its
    // buckets before the fine transport starts) with a payload counter of their own in ctl.
    Bkt bkt(int fine) const {
        Bkt b;
        b.bid = bbid.p; b.n = bn.p; b.sref = bsref.p; b.pool = bpool.p;
        b.pcount = ctl.p + C_BPC + (fine ? NSH * SH_STRIDE : 0);
        b.pcap = pcap;
        b.ftag = bftag.p;
        b.flist = bflist.p;
        return b;
    }
    Tally tally() const {
        Tally t;
        t.edep = edep.p; t.w_cut = w_cut.p; t.E_cut = E_cut.p; t.w_nuc = w_nuc.p; t.E_nuc = E_nuc.p;
        t.w_bnd = w_bnd.p; t.E_bnd = E_bnd.p;
        return t;
    }
    uint32_t* nocc(int k) const { return ctl.p + (k ? C_NOCC1 : C_NOCC0); }
};

namespace {

// Launch sizes follow the list-length hints rounded up to powers of two, so that captured chunks
// of iterations can be replayed while the lists keep growing or shrinking.
uint32_t pow2_at_least(uint32_t x) {
    uint32_t p = 1;
    while (p < x && p < (1u << 31)) p <<= 1;
    return p;
}

int blocks_for(uint64_t items, int per_block, int cap) {
    uint64_t b = (items + per_block - 1) / per_block;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(b, (uint64_t)cap));
}

int b_threshold(const sm2d_ctx* c, float E) {  // compute_b_E_threshold (k1k6_pipeline.cuh:419-422)
    const std::vector<float>& e = c->h_ee;
    int NE = c->g.N_E, bin;
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
    return bin / c->g.n_e_local;
}

Tp make_tp(const sm2d_ctx* c, const sm2d_transport* t) {
    Tp p;
    std::memset(&p, 0, sizeof(p));  // padding too: the bytes key the chunk graphs
    p.E_on = t->E_fine_on;
    p.E_off = t->E_fine_off;
    p.wmin = t->weight_active_min;
    p.step_coarse = t->step_coarse;
    p.b_on = b_threshold(c, t->E_fine_on);
    p.b_off = b_threshold(c, t->E_fine_off);
    p.rule = t->step_rule;
    p.strag = t->enable_straggling;
    p.nuc = t->enable_nuclear;
    p.surr = (c->surr_on && c->surr.p) ? (c->surr_on == 2 && c->mlp_wb ? 2 : 1) : 0;
    if (!c->h_R.empty())
        p.R_on = lookup_R_tab(c->h_ee.data(), c->h_R.data(), c->h_lR.data(), c->h_lE.data(), c->g.N_E,
                              std::max(p.E_on, c->h_ee[0]));
    return p;
}

__global__ void k_ctl_reset(uint32_t* ctl) {
    ctl[C_DONE] = 0;
    ctl[C_DRIFT] = 0;
    ctl[C_IT] = 0;
    ctl[C_DONE_IT] = 0;
}

__global__ void k_undo_open(uint32_t* ctl) {
    if (!ctl[C_DONE]) {
        ctl[C_EP] -= 1;
        ctl[C_IT] -= 1;
    }
}

// Lists up to these lengths are built by one workgroup (one launch instead of three).
constexpr uint32_t SMALL_LIST = 8192, SMALL_WORDS = 16384, SMALL_AUDIT = 8192;

// Tiles of 256 listed cells per phase (k_acc_tiles).
inline uint32_t acc_tcap(const sm2d_ctx* c) { return c->lcap / 256 + 2; }

// Occupancy bitmap -> ascending occupied-cell list of buffer k.
void build_occ(sm2d_ctx* c, int k) {
    if (c->nwords <= SMALL_WORDS) {
        k_occ_small<<<1, 1024, 0, c->stream>>>(c->bitmap.p, c->nwords, c->occ[k].p, c->ocap, c->nocc(k), c->ctl.p,
                                               c->flags.p);
        return;
    }
    const int nt = blocks_for(c->nwords, 256, 1 << 20);
    k_occ_count<<<nt, 256, 0, c->stream>>>(c->bitmap.p, c->nwords, c->tcount2.p, c->ctl.p);
    // n_items = nwords is held in ctl[15]
    k_scan<1><<<1, 1024, 0, c->stream>>>(c->tcount2.p, c->ctl.p + 15, 256, c->ctl.p, c->nocc(k), c->ocap,
                                         c->flags.p, 0u);
    k_occ_emit<<<nt, 256, 0, c->stream>>>(c->bitmap.p, c->nwords, c->tcount2.p, c->occ[k].p, c->ocap, c->ctl.p);
}

// Compaction targets of buffer `in`'s occupied list (occ = null: no compaction).
ListOut list_out(sm2d_ctx* c, int in) {
    ListOut L;
    L.occ = c->occ[in].p;
    L.flag = c->flag.p;
    L.act = c->act.p;
    L.crs = c->crs.p;
    L.ltag = c->ltag.p;
    L.lpos = c->lpos.p;
    L.bitmap = c->bitmap.p;
    L.flags = c->flags.p;
    L.cap = c->lcap;
    L.skip = 0;
    return L;
}

// K1 + K2 on psi_in (buffer `in`): mask, flag counts, ordered compaction.  mask = false: the
// previous iteration's closing audit already decided the mask and flags (k_audit_cells);
// compact = false: it also built the lists (k_audit_final).
template <class LY>
void launch_mask_compact(sm2d_ctx* c, const Tp& P, int in, int grid_entries, bool mask = true, bool compact = true,
                         uint32_t skip = 0) {
    const int ge = blocks_for(grid_entries, 1, 1 << 16);
    const int gt = blocks_for(grid_entries, 256, 4096);
    if (mask)
        k_mask<LY><<<ge, 64, 0, c->stream>>>(c->tabs(), P, c->psi(in), c->occ[in].p, c->nocc(in), c->maskv.p,
                                             c->ctl.p, c->flag.p, c->win.p, c->ein.p);
    if (!compact) return;
    if ((uint32_t)grid_entries <= SMALL_LIST) {
        ListOut L = list_out(c, in);
        L.skip = skip;
        k_compact_small<<<1, 1024, 0, c->stream>>>(L, c->nocc(in), c->ctl.p);
        return;
    }
    k_flag_count<<<gt, 256, 0, c->stream>>>(c->flag.p, c->nocc(in), c->tcount.p, c->ctl.p);
    k_scan<2><<<1, 1024, 0, c->stream>>>(c->tcount.p, c->nocc(in), 256, c->ctl.p, nullptr, c->lcap, c->flags.p,
                                          skip);
    k_emit<<<gt, 256, 0, c->stream>>>(c->occ[in].p, c->nocc(in), c->flag.p, c->tcount.p, c->act.p, c->crs.p,
                                      c->ltag.p, c->lpos.p, c->bitmap.p, c->ctl.p);
}

// K5 stage 1 over (nin entries of K1 sums, nout occupied cells of buffer k).  Returns the
// per-cell sums for k_audit_final to tile itself (short lists), else null after the tile pass.
// With P, buffer k's K1 decision for the next iteration is made in the same pass.
template <class LY>
// tiles = false: the caller launches the entry tiles itself (with the accumulator tiles, k_tiles).
const double* launch_audit_cells(sm2d_ctx* c, int k, const uint32_t* nin, bool k4, int force, int grid_entries,
                                 const Tp* P = nullptr, bool tiles = true) {
    const int ge = blocks_for(grid_entries, 1, 1 << 16);
    const int gt = blocks_for(grid_entries, 256, 4096);
    k_audit_cells<LY><<<ge, 64, 0, c->stream>>>(c->tabs(), c->psi(k), nin, c->win.p, c->ein.p, c->occ[k].p,
                                                c->nocc(k), k4 ? c->k4w.p : nullptr, k4 ? c->k4e.p : nullptr,
                                                c->pe.p, c->ctl.p, force, P ? *P : Tp{}, P ? c->maskv.p : nullptr,
                                                c->flag.p);
    if ((uint32_t)grid_entries <= SMALL_AUDIT) return c->pe.p;
    if (tiles) k_audit_tiles<<<gt, 256, 0, c->stream>>>(c->pe.p, nin, c->nocc(k), c->part.p, c->ctl.p, force);
    return nullptr;
}

template <class LY>
// opener: open the iteration here (first of a run); open_next: the closing audit kernel opens
// the next iteration and clears its output buffer, which is this iteration's input.
// compacted: the previous iteration's closing kernel built this iteration's lists; compact_next:
// this iteration's closing kernel builds the next one's (short lists only).  skip: phases left
// out (bit 0 coarse, bit 1 fine; short lists only): their lists are empty in every iteration that
// runs, since lists that need them pause the transport first.
void launch_iteration(sm2d_ctx* c, const Tp& P, bool opener, bool open_next, bool compacted = false,
                      bool compact_next = false, uint32_t skip = 0) {
    const int in = c->in, out = 1 - c->in;
    Tabs T = c->tabs();
    const uint32_t h = pow2_at_least(std::max<uint32_t>(c->hint, 64));
    const uint32_t hocc = pow2_at_least(c->hint_occ);
    const int gm = blocks_for(4ull * h, 256, 4096);  // occupied-list tiles
    if (opener) k_clear<<<gm, 256, 0, c->stream>>>(c->psi(out), c->occ[out].p, c->nocc(out), c->ctl.p);
    c->mark(0);
    // an iteration opened by the previous one's closing kernel has its K1 decision from there
    launch_mask_compact<LY>(c, P, in, (int)hocc, opener, !compacted, skip);
    c->mark(1);
    for (int fine = 0; fine < 2; ++fine) {
        if (skip & (1u << fine)) continue;
        const uint32_t* list = fine ? c->act.p : c->crs.p;
        const uint32_t* nl = c->ctl.p + (fine ? C_NACT : C_NCRS);
        double* acc = fine ? c->acc_f.p : c->acc_c.p;
        const int e0 = fine ? 6 : 2;
        const int tw = Tuner::tw(c->cand[fine]);
        // launch sizes from the longer list's hint, for both phases: a list can grow within a
        // chunk (components slowing into the fine phase), and a grid sized from the last chunk's
        // short fine list would stride over the grown list with too few workgroups; surplus
        // workgroups return at once.  A phase whose list was empty is left out of the chunk instead
        // (skip).
        const int gb = blocks_for(h, 1, 65536);
        // timed launches take their events from the dispatch itself (the kernel's own start and
        // end), so stage times and the tuner see the kernel without marker packets around it
        hipEvent_t e_start = c->rec ? c->ev[(size_t)c->ev_it * sm2d_ctx::EPI + e0] : nullptr;
        hipEvent_t e_stop = c->rec ? c->ev[(size_t)c->ev_it * sm2d_ctx::EPI + e0 + 1] : nullptr;
#define SM2D_TRANSPORT(F, W, I)                                                                               \
    do {                                                                                                      \
        if (c->rec)                                                                                           \
            hipExtLaunchKernelGGL(k_transport<LY, F, W, I>, dim3(gb), dim3(64 * W),                           \
                                  (uint32_t)transport_dyn_lds(c->g.N_E, W), c->stream,                        \
                                  e_start, e_stop, 0u, T, P, c->psi(in), c->psi(out), c->bkt(fine), list, nl, \
                                  (const uint32_t*)c->ltag.p, (const uint32_t*)c->lpos.p, c->ctl.p,           \
                                  c->tally(), acc,                                                            \
                                  c->flags.p);                                                                \
        else                                                                                                  \
            k_transport<LY, F, W, I><<<gb, 64 * W, transport_dyn_lds(c->g.N_E, W), c->stream>>>(                 \
                T, P, c->psi(in), c->psi(out), c->bkt(fine),                                                      \
                                                                 list, nl, c->ltag.p, c->lpos.p, c->ctl.p,    \
                                                                 c->tally(),                                  \
                                                                 acc, c->flags.p);                            \
    } while (0)
        // learned-physics mode 2 (the network inside the fine transport): its own one-wave instances
        const bool infer = fine && P.surr == 2;
        if (infer) SM2D_TRANSPORT(true, 1, true);
        else if (fine) {
            if (tw == 1) SM2D_TRANSPORT(true, 1, false);
            else if (tw == 2) SM2D_TRANSPORT(true, 2, false);
            else SM2D_TRANSPORT(true, 4, false);
        } else {
            if (tw == 1) SM2D_TRANSPORT(false, 1, false);
            else if (tw == 2) SM2D_TRANSPORT(false, 2, false);
            else SM2D_TRANSPORT(false, 4, false);
        }
#undef SM2D_TRANSPORT
        if (infer) k_foreign<LY, true, true><<<1, 64, 0, c->stream>>>(T, P, c->psi(in), c->bkt(fine), list,
                                                                    c->ltag.p, c->ctl.p, c->flags.p);
        else if (fine) k_foreign<LY, true, false><<<1, 64, 0, c->stream>>>(T, P, c->psi(in), c->bkt(fine), list,
                                                                         c->ltag.p, c->ctl.p, c->flags.p);
        else k_foreign<LY, false, false><<<1, 64, 0, c->stream>>>(T, P, c->psi(in), c->bkt(fine), list, c->ltag.p,
                                                                  c->ctl.p, c->flags.p);
        uint32_t* seg = c->ctl.p + C_DSEG + (fine ? NSH * SH_STRIDE : 0);
        const uint32_t seg_cap = c->dcap / NSH;
        k_dest<<<blocks_for(4ull * h, 256, 4096), 256, 0, c->stream>>>(list, nl, c->g.Nx, c->g.Nz, c->Ns, c->bn.p,
                                                                       c->dtag.p, c->ctl.p, fine, c->dlist.p, seg,
                                                                       seg_cap, c->bitmap.p, c->flags.p);
        uint32_t* moved = c->prof ? c->gmoved.p + (size_t)(2 * c->ev_it + fine) * NSH * SH_STRIDE : nullptr;
        const unsigned gg = (blocks_for(4ull * h, 1, 65536) + NSH - 1) / NSH * NSH;
        const int kib = Tuner::kib(c->cand[2 + fine]);
#define SM2D_GATHER(K)                                                                                        \
    do {                                                                                                      \
        if (c->rec)                                                                                           \
            hipExtLaunchKernelGGL(c->prof ? k_gather<LY, true, K> : k_gather<LY, false, K>, dim3(gg), dim3(64), \
                                  0u, c->stream, c->ev[(size_t)c->ev_it * sm2d_ctx::EPI + e0 + 2],            \
                                  c->ev[(size_t)c->ev_it * sm2d_ctx::EPI + e0 + 3], 0u, T, c->psi(out),      \
                                  c->bkt(fine), (const uint32_t*)c->dlist.p, (const uint32_t*)seg, seg_cap,   \
                                  (const uint32_t*)c->ltag.p, (const uint32_t*)c->lpos.p,                     \
                                  (const uint32_t*)c->ctl.p, fine, c->k4w.p, c->k4e.p, c->flags.p, moved);    \
        else                                                                                                  \
            k_gather<LY, false, K><<<gg, 64, 0, c->stream>>>(T, c->psi(out), c->bkt(fine), c->dlist.p, seg,      \
                                                             seg_cap, c->ltag.p, c->lpos.p, c->ctl.p, fine,     \
                                                             c->k4w.p, c->k4e.p, c->flags.p, moved);            \
    } while (0)
        if (kib == 16) SM2D_GATHER(16);
        else if (kib == 8) SM2D_GATHER(8);
        else SM2D_GATHER(4);
#undef SM2D_GATHER
    }
    build_occ(c, out);
    c->mark(10);
    // long lists: the listed cells' accumulators in tiles too, so the closing workgroup reads tile
    // partials instead of every cell's; with a long occupied list as well, one launch does the
    // entry tiles and the accumulator tiles (k_tiles)
    const bool acc_tiles = h > SMALL_AUDIT, audit_tiles = hocc > SMALL_AUDIT;
    const double* pe = launch_audit_cells<LY>(c, out, c->nocc(in), true, 0, (int)hocc, open_next ? &P : nullptr,
                                              !acc_tiles);
    ListOut L = list_out(c, out);
    L.skip = skip;
    if (!(open_next && compact_next)) L.occ = nullptr;
    if (acc_tiles) {
        const uint32_t ga = audit_tiles ? (uint32_t)blocks_for(hocc, 256, 4096) : 0u;
        k_tiles<<<ga + blocks_for(h, 256, 4096), 256, 0, c->stream>>>(
            c->pe.p, c->nocc(in), c->nocc(out), c->part.p, ga, c->acc_c.p, c->ctl.p + C_NCRS, c->acc_f.p,
            c->ctl.p + C_NACT, c->accpart.p, acc_tcap(c), c->ctl.p);
    }
    k_audit_final<<<open_next ? 1 + gm : 1, 256, 0, c->stream>>>(
        c->part.p, c->nocc(in), c->nocc(out), c->acc_c.p, c->ctl.p + C_NCRS, c->acc_f.p, c->ctl.p + C_NACT, nullptr,
        c->hist.p, HROWS, c->ctl.p, 0, pe, c->psi(in), c->occ[in].p, c->nocc(in), L, P.surr,
        acc_tiles ? c->accpart.p : nullptr, acc_tcap(c));
    c->mark(11);
    c->in = out;  // K6
}

// Totals (W, E) of buffer k over its occupied list, after the transport (forced pass).
template <class LY>
int buffer_totals(sm2d_ctx* c, int k, double& W, double& E) {
    const double* pe = launch_audit_cells<LY>(c, k, nullptr, false, 1, (int)std::max<uint32_t>(c->hint_occ, 1024));
    ListOut none{};
    k_audit_final<<<1, 256, 0, c->stream>>>(c->part.p, nullptr, c->nocc(k), nullptr, nullptr, nullptr, nullptr,
                                            c->k5v.p, nullptr, 1, c->ctl.p, 1, pe, Psi{}, nullptr, nullptr, none,
                                            0, nullptr, 0u);
    HIPCHK(hipGetLastError());
    double v[NV];
    HIPCHK(hipMemcpyAsync(v, c->k5v.p, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    W = v[R_WOUT];
    E = v[R_EOUT];
    return SM2D_OK;
}

// Stage times of the first n_valid iterations of the chunk just synchronised (iterations queued
// after the transport ended return at once and are not counted).
void collect_timing(sm2d_ctx* c, int n_valid) {
    if (!c->prof) return;
    uint32_t moved[2 * CHUNK] = {};
    if (n_valid > 0) {
        const size_t per = (size_t)NSH * SH_STRIDE;
        std::vector<uint32_t> sh((size_t)2 * std::min(n_valid, CHUNK) * per);
        if (hipMemcpy(sh.data(), c->gmoved.p, sizeof(uint32_t) * sh.size(), hipMemcpyDeviceToHost) != hipSuccess)
            return;
        for (size_t k = 0; k < sh.size() / per; ++k)
            for (int s = 0; s < NSH; ++s) moved[k] += sh[k * per + (size_t)s * SH_STRIDE];
    }
    for (int r = 0; r < n_valid; ++r) {
        hipEvent_t* e = c->ev.data() + (size_t)r * sm2d_ctx::EPI;
        auto span = [&](int a, int b) -> double {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e[a], e[b]);
            return (double)ms;
        };
        c->timing.mask_ms += span(0, 1);
        c->timing.transport_ms += span(2, 3) + span(6, 7);
        c->timing.coarse_transport_ms += span(2, 3);
        c->timing.fine_transport_ms += span(6, 7);
        c->timing.gather_ms += span(4, 5) + span(8, 9);
        c->timing.coarse_gather_ms += span(4, 5);
        c->timing.fine_gather_ms += span(8, 9);
        c->timing.coarse_gather_values += moved[2 * r];
        c->timing.fine_gather_values += moved[2 * r + 1];
        c->timing.audit_ms += span(9, 11);
        c->timing.transport_launches += 2;
        c->timing.gather_launches += 2;
    }
}

int check_flags(sm2d_ctx* c, const uint32_t* f) {
    if (f[F_PSI_FULL]) return err("phase-space payload pool full (raise pool capacity)", SM2D_ERR_CAPACITY);
    if (f[F_BKT_FULL]) return err("bucket payload pool full (raise pool capacity)", SM2D_ERR_CAPACITY);
    if (f[F_DEST_FULL]) return err("destination / list capacity exceeded", SM2D_ERR_CAPACITY);
    return SM2D_OK;
}

void fill_report(const double* v, double Wso, double Wss, double Eso, double Ess, int src, sm2d_k5_report* r) {
    std::memset(r, 0, sizeof(*r));
    r->struct_size = sizeof(*r);
    r->W_in = v[R_WIN];
    r->E_in = v[R_EIN];
    r->W_out = v[R_WOUT];
    r->E_out = v[R_EOUT];
    const double* a = v + NR1;
    r->W_cutoff = a[A_WCUT];
    r->W_nuclear = a[A_WNUC];
    r->W_boundary = a[A_WBND];
    r->E_dep = a[A_EDEP];
    r->E_cutoff = a[A_ECUT];
    r->E_nuclear = a[A_ENUC];
    r->E_boundary = a[A_EBND];
    r->W_transport_drop = std::max(0.0, a[A_DROPW] + v[R_K4W]);
    r->E_transport_drop = std::max(0.0, a[A_DROPE] + v[R_K4E]);
    if (src) {
        double x = std::max(0.0, Wso), y = std::max(0.0, Wss), e1 = std::max(0.0, Eso), e2 = std::max(0.0, Ess);
        r->W_in += x + y;
        r->E_in += e1 + e2;
        r->W_source_out_of_grid = x;
        r->W_source_slot_drop = y;
        r->E_source_out_of_grid = e1;
        r->E_source_slot_drop = e2;
    }
    double Wr = r->W_out + r->W_cutoff + r->W_nuclear + r->W_boundary + r->W_transport_drop +
                r->W_source_out_of_grid + r->W_source_slot_drop;
    r->W_error = std::fabs(r->W_in - Wr) / std::max(r->W_in, 1e-20);
    r->W_pass = (std::isfinite(r->W_error) && r->W_error < 1e-6) ? 1 : 0;
    double Er = r->E_out + r->E_dep + r->E_cutoff + r->E_nuclear + r->E_boundary + r->E_transport_drop +
                r->E_source_out_of_grid + r->E_source_slot_drop;
    r->E_error = std::fabs(r->E_in - Er) / std::max(r->E_in, 1e-20);
    r->E_pass = (std::isfinite(r->E_error) && r->E_error < 1e-5) ? 1 : 0;
}

template <class T>
void key_put(std::vector<uint8_t>& k, const T& v) {
    const uint8_t* b = reinterpret_cast<const uint8_t*>(&v);
    k.insert(k.end(), b, b + sizeof(T));
}

// Queues R iterations as one graph launch.  A chunk is captured once per key (iteration count,
// buffer parity, opening mode, transport shapes, launch sizes, physics switches, surrogate table)
// and replayed afterwards; every launch of the chunk reads its lists, counts and epochs from the
// device control block, so a replay is the same work as the eager launches.
template <class LY>
int launch_chunk_graph(sm2d_ctx* c, const Tp& P, int R, bool first_opens, bool fused_open, uint32_t skip = 0) {
    std::vector<uint8_t> key;
    key_put(key, R);
    key_put(key, skip);
    key_put(key, c->in);
    key_put(key, (int)first_opens);
    key_put(key, (int)fused_open);
    for (int kind = 0; kind < Tuner::NK; ++kind) key_put(key, c->cand[kind]);
    key_put(key, pow2_at_least(std::max<uint32_t>(c->hint, 64)));
    key_put(key, pow2_at_least(c->hint_occ));
    key_put(key, pow2_at_least(std::max<uint32_t>(c->hint_ph[0], 64)));
    key_put(key, pow2_at_least(std::max<uint32_t>(c->hint_ph[1], 64)));
    key_put(key, P);
    key_put(key, (const void*)c->surr.p);
    key_put(key, (const void*)c->rowmat.p);
    key_put(key, c->surr_nodes);
    ++c->graph_clock;
    for (sm2d_ctx::ChunkGraph& g : c->graphs)
        if (g.key == key) {
            g.used = c->graph_clock;
            HIPCHK(hipGraphLaunch(g.exec, c->stream));
            if (R & 1) c->in = 1 - c->in;
            ++c->graph_replays;
            return SM2D_OK;
        }
    const int in0 = c->in;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    // inside a chunk, each iteration's closing kernel builds the next one's short lists
    const bool fold = fused_open && pow2_at_least(c->hint_occ) <= SMALL_LIST;
    for (int r = 0; r < R; ++r)
        launch_iteration<LY>(c, P, first_opens && r == 0, fused_open, fold && r > 0, fold && r + 1 < R, skip);
    hipGraph_t graph = nullptr;
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
    if (le != hipSuccess || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        c->in = in0;
        return err(std::string("chunk capture: ") + hipGetErrorString(le != hipSuccess ? le : ce), SM2D_ERR_HIP);
    }
    sm2d_ctx::ChunkGraph g;
    g.key = std::move(key);
    g.used = c->graph_clock;
    const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        c->in = in0;
        return err(std::string("chunk graph instantiate: ") + hipGetErrorString(ie), SM2D_ERR_HIP);
    }
    HIPCHK(hipGraphLaunch(g.exec, c->stream));
    ++c->graph_captures;
    constexpr size_t MAX_GRAPHS = 48;
    if (c->graphs.size() >= MAX_GRAPHS) {  // drop the least recently used
        size_t lru = 0;
        for (size_t i = 1; i < c->graphs.size(); ++i)
            if (c->graphs[i].used < c->graphs[lru].used) lru = i;
        (void)hipGraphExecDestroy(c->graphs[lru].exec);
        c->graphs.erase(c->graphs.begin() + (long)lru);
    }
    c->graphs.push_back(std::move(g));
    return SM2D_OK;
}

template <class LY>
int run_iters(sm2d_ctx* c, const sm2d_transport* tp, int n_iter, int src_terms, sm2d_k5_report* last,
              sm2d_audit* au) {
    Tp P = make_tp(c, tp);
    if (au) au->surrogate_active = P.surr;
    k_ctl_reset<<<1, 1, 0, c->stream>>>(c->ctl.p);
    HIPCHK(hipGetLastError());
    const int in0 = c->in;
    const int chunk = tp->fail_fast_on_audit ? 1 : CHUNK;
    const bool fused_open = !tp->fail_fast_on_audit;
    const int opened_ahead = fused_open ? 1 : 0;  // ctl[C_IT] counts the next, already opened, iteration
    int queued = 0, audited = 0;
    // phases whose lists were empty when the last chunk ended; a graph chunk leaves their kernels
    // out (short lists) until a list needs them, which pauses the transport (once per run)
    uint32_t last_nact = 1, last_ncrs = 1;
    bool paused_once = false;
    double Wso = 0, Wss = 0, Eso = 0, Ess = 0;
    for (int s = 0; s < c->S; ++s) {
        Wso += c->src[s * 7 + 1];
        Wss += c->src[s * 7 + 2];
        Eso += c->src[s * 7 + 4];
        Ess += c->src[s * 7 + 5];
    }
    std::vector<double> rows((size_t)HROWS * NV);
    while (queued < n_iter) {
        const int R = std::min(chunk, n_iter - queued);
        // transport shapes of this chunk (see Tuner)
        int bkt[2];
        bool explore[2];
        for (int ph = 0; ph < 2; ++ph) {
            bkt[ph] = Tuner::bucket(c->hint_ph[ph] / 2);
            explore[ph] = !c->tune.forced && bkt[ph] > 0 && !c->tune.done[ph][bkt[ph]];
        }
        c->rec = c->prof || explore[0] || explore[1];
        if (c->prof)
            HIPCHK(hipMemsetAsync(c->gmoved.p, 0, sizeof(uint32_t) * 2 * CHUNK * NSH * SH_STRIDE, c->stream));
        c->cand_log.assign((size_t)Tuner::NK * R, 0);
        // shape of tuned launch `kind` (phase kind & 1) in chunk iteration r
        auto shape = [&](int kind, int r) {
            const int ph = kind & 1;
            if (kind >= 2 && c->tune.forced_kib)
                return c->tune.forced_kib == 16 ? 0 : c->tune.forced_kib == 8 ? 1 : 2;
            if (c->tune.forced) return kind < 2 ? (c->tune.forced == 1 ? 0 : c->tune.forced == 2 ? 1 : 2) : 0;
            if (bkt[ph] == 0) return kind < 2 ? 0 : Tuner::dflt(kind);
            if (kind < 2 && bkt[ph] >= Tuner::ONE_WAVE) return 0;  // long lists: one wave per cell
            if (explore[ph]) return r % Tuner::NC;
            return c->tune.best(kind, bkt[ph]);
        };
        if (!c->rec && c->graphs_on) {
            // the same shapes for every iteration of the chunk (no exploration)
            for (int kind = 0; kind < Tuner::NK; ++kind) {
                const int cd = shape(kind, 0);
                c->cand[kind] = cd;
                for (int r = 0; r < R; ++r) c->cand_log[(size_t)Tuner::NK * r + kind] = cd;
            }
            uint32_t skip = 0;
            if (fused_open && queued > 0 && !paused_once) {
                if (last_ncrs == 0) skip |= 1u;
                if (last_nact == 0) skip |= 2u;
                if (skip == 3u) skip = 0;
            }
            if (skip) c->skip_chunks++;
            const int rc = launch_chunk_graph<LY>(c, P, R, fused_open ? queued == 0 : true, fused_open, skip);
            if (rc) return rc;
        } else for (int r = 0; r < R; ++r) {
            for (int kind = 0; kind < Tuner::NK; ++kind) {
                const int cd = shape(kind, r);
                c->cand[kind] = cd;
                c->cand_log[(size_t)Tuner::NK * r + kind] = cd;
            }
            c->ev_it = r;
            // fail-fast keeps psi_in of a failing iteration, so there every iteration opens itself
            const bool fold = fused_open && pow2_at_least(c->hint_occ) <= SMALL_LIST;
            launch_iteration<LY>(c, P, fused_open ? (queued == 0 && r == 0) : true, fused_open, fold && r > 0,
                                 fold && r + 1 < R);
        }
        const int first_it = queued;  // global index of the chunk's first iteration
        queued += R;
        HIPCHK(hipGetLastError());
        uint32_t ctl[NCTL], f[NFLAG];
        HIPCHK(hipMemcpyAsync(ctl, c->ctl.p, sizeof(ctl), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(rows.data(), c->hist.p, sizeof(double) * rows.size(), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        int rc = check_flags(c, f);
        if (rc) return rc;
        const bool done = ctl[C_DONE] == 1, paused = ctl[C_DONE] == 2;
        const int upto = done ? (int)ctl[C_DONE_IT] - 1 : (int)ctl[C_IT] - opened_ahead;  // audited iterations
        uint32_t maxlist = 0;
        for (int it = audited; it < upto; ++it) {
            const double* v = rows.data() + (size_t)(it % HROWS) * NV;
            sm2d_k5_report r;
            fill_report(v, Wso, Wss, Eso, Ess, (src_terms && it == 0) ? 1 : 0, &r);
            if (last) *last = r;
            if (au) {
                const double* a = v + NR1;
                au->W_transport_drop += a[A_DROPW] + v[R_K4W];
                au->E_transport_drop += a[A_DROPE] + v[R_K4E];
                double Er = r.E_out + r.E_dep + r.E_cutoff + r.E_nuclear + r.E_boundary + r.E_transport_drop +
                            r.E_source_out_of_grid + r.E_source_slot_drop;
                double res = r.E_in - Er;
                if (std::isfinite(res)) au->E_audit_residual += res;
                au->max_W_error = std::max(au->max_W_error, r.W_error);
                au->max_E_error = std::max(au->max_E_error, r.E_error);
                if (!r.W_pass) au->W_fail_iters++;
                if (!r.E_pass) au->E_fail_iters++;
                au->max_cell_W_closure = std::max(au->max_cell_W_closure, v[NR1 + NACC]);
                au->max_cell_E_closure = std::max(au->max_cell_E_closure, v[NR1 + NACC + 1]);
                au->cell_bin_updates += (uint64_t)a[A_UPD];
                au->fine_cell_bin_updates += (uint64_t)v[IDX_FUPD];
                au->surrogate_substeps += (uint64_t)a[A_SURR];
                au->processed_cells += (uint64_t)v[IDX_LIST];
                au->processed_components += (uint64_t)a[A_COMPS];
            }
            if (c->prof) {
                c->timing.cell_bin_updates += (uint64_t)v[NR1 + A_UPD];
                c->timing.fine_updates += (uint64_t)v[IDX_FUPD];
                c->timing.coarse_updates += (uint64_t)v[NR1 + A_UPD] - (uint64_t)v[IDX_FUPD];
            }
            if (tp->fail_fast_on_audit && (!r.W_pass || !r.E_pass)) {
                c->in = in0 ^ (it & 1);  // psi_in of the failing iteration
                return err("K5 audit fail-fast", SM2D_ERR_AUDIT);
            }
        }
        collect_timing(c, upto - audited);
        for (int ph = 0; ph < 2; ++ph) {
            if (!explore[ph]) continue;
            const int e0 = ph ? 6 : 2;
            bool all = true;
            for (int kind = ph; kind < Tuner::NK; kind += 2) {  // the phase's transport, then its gather
                const int e = e0 + (kind < 2 ? 0 : 2);
                for (int it = std::max(audited, first_it); it < upto; ++it) {
                    const int r = it - first_it;
                    float ms = 0.0f;
                    if (hipEventElapsedTime(&ms, c->ev[(size_t)r * sm2d_ctx::EPI + e],
                                            c->ev[(size_t)r * sm2d_ctx::EPI + e + 1]) != hipSuccess)
                        continue;
                    const int cd = c->cand_log[(size_t)Tuner::NK * r + kind];
                    c->tune.ms[kind][bkt[ph]][cd] += ms;
                    c->tune.cnt[kind][bkt[ph]][cd] += 1;
                }
                if (kind < 2 && bkt[ph] >= Tuner::ONE_WAVE) continue;  // not explored (best() is one wave)
                for (int cd = 0; cd < Tuner::NC; ++cd) all = all && c->tune.cnt[kind][bkt[ph]][cd] >= 2;
            }
            if (all) c->tune.done[ph][bkt[ph]] = true;
        }
        audited = upto;
        // K5 of an iteration that flagged the learned physics switched the following ones to the
        // plain physics on the device (the control word C_DRIFT)
        if (P.surr && ctl[C_DRIFT] && au) au->surrogate_fallback = 1;
        maxlist = std::max(ctl[C_NACT], ctl[C_NCRS]);
        c->hint = std::max<uint32_t>(64, 2 * maxlist);
        c->hint_ph[0] = std::max<uint32_t>(2, 2 * ctl[C_NCRS]);
        c->hint_ph[1] = std::max<uint32_t>(2, 2 * ctl[C_NACT]);
        c->hint_occ = std::max<uint32_t>(64, 2 * std::max(ctl[C_NOCC0], ctl[C_NOCC1]));
        if (au) au->iterations = done ? (int)ctl[C_DONE_IT] : (int)ctl[C_IT] - opened_ahead;
        last_nact = ctl[C_NACT];
        last_ncrs = ctl[C_NCRS];
        if (paused) {  // the iterations from `upto` on did not run: queue them again with both phases
            const uint32_t zero = 0;
            HIPCHK(hipMemcpyAsync(c->ctl.p + C_DONE, &zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            queued = upto;
            c->in = in0 ^ (upto & 1);
            paused_once = true;
            c->pauses++;
        }
        if (done) {
            if (au) au->terminated = 1;
            c->in = in0 ^ (int)((ctl[C_DONE_IT] - 1u) & 1u);  // the buffer of the terminating iteration
            break;
        }
    }
    if (c->prof) c->timing.iterations += (uint64_t)audited;
    // A run that stopped at its iteration count has opened one more iteration: take it back, so
    // that the epochs (and the mask hysteresis that follows them) continue in the next run.
    if (fused_open) k_undo_open<<<1, 1, 0, c->stream>>>(c->ctl.p);
    HIPCHK(hipGetLastError());
    return SM2D_OK;
}

// A new phase space (injected or loaded) with `no` occupied cells: the first chunk's launches are
// sized from it, not from the lists the previous run ended with.
void seed_hints(sm2d_ctx* c, uint32_t no) {
    const uint32_t h = (uint32_t)std::min<size_t>(c->lcap, std::max<size_t>(64, 2 * (size_t)no));
    c->hint_occ = std::max<uint32_t>(64, 2 * no);
    c->hint = c->hint_ph[0] = c->hint_ph[1] = h;
}

template <class LY>
int inject_all(sm2d_ctx* c, const sm2d_spot* sp, int n) {
    int nmax = 1;
    for (int s = 0; s < n; ++s) {
        const bool gauss = (sp[s].n_samples > 1) || (sp[s].sigma_x > 0) || (sp[s].sigma_theta > 0);
        if (gauss && sp[s].n_samples <= 0) return err("n_samples must be positive");
        nmax = std::max(nmax, gauss ? sp[s].n_samples : 1);
    }
    if (c->samples.n < (size_t)n * nmax) HIPCHK(c->samples.alloc((size_t)n * nmax));
    HIPCHK(hipMemcpyAsync(c->spots.p, sp, sizeof(sm2d_spot) * n, hipMemcpyHostToDevice, c->stream));
    Tabs T = c->tabs();
    k_inject_samples<LY><<<blocks_for((uint64_t)n * nmax, 256, 8192), 256, 0, c->stream>>>(T, c->spots.p, n, nmax,
                                                                                           c->samples.p);
    k_inject_apply<LY><<<n, 64, 0, c->stream>>>(c->psi(c->in), c->spots.p, c->samples.p, nmax, c->injacc.p,
                                                c->bitmap.p, c->flags.p);
    k_ctl_reset<<<1, 1, 0, c->stream>>>(c->ctl.p);  // a finished transport must not gate the list build
    build_occ(c, c->in);
    HIPCHK(hipGetLastError());
    std::vector<double> acc((size_t)n * 6);
    uint32_t f[NFLAG], no = 0;
    HIPCHK(hipMemcpyAsync(acc.data(), c->injacc.p, sizeof(double) * acc.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&no, c->nocc(c->in), sizeof(no), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    seed_hints(c, no);
    int rc = check_flags(c, f);
    if (rc) return rc;
    // represented energy of psi_in (gpu_transport_wrapper.cu:125-180,443-457); per batch
    double W, E;
    rc = buffer_totals<LY>(c, c->in, W, E);
    if (rc) return rc;
    double Ein = 0.0;
    for (int s = 0; s < n; ++s) {
        for (int k = 0; k < 6; ++k) c->src[s * 7 + k] = acc[s * 6 + k];
        Ein += acc[s * 6 + 3];
    }
    double loss = Ein - E;
    if (std::fabs(loss) < 1e-12) loss = 0.0;
    for (int s = 0; s < c->S; ++s) c->src[s * 7 + 6] = 0.0;
    c->src[6] = loss;  // the batch's representation loss is booked on spot 0
    return SM2D_OK;
}

int zero_state(sm2d_ctx* c) {
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipMemsetAsync(c->bid[k].p, 0xFF, sizeof(uint32_t) * c->bid[k].n, c->stream));
        HIPCHK(hipMemsetAsync(c->ns[k].p, 0, sizeof(uint32_t) * c->ns[k].n, c->stream));
        HIPCHK(hipMemsetAsync(c->pcount[k].p, 0, sizeof(uint32_t) * NSH * SH_STRIDE, c->stream));
    }
    const size_t N = c->N;
    for (DBuf<double>* b : {&c->edep, &c->E_cut, &c->E_nuc, &c->E_bnd, &c->k4w, &c->k4e})
        HIPCHK(hipMemsetAsync(b->p, 0, sizeof(double) * N, c->stream));
    for (DBuf<float>* b : {&c->w_cut, &c->w_nuc, &c->w_bnd}) HIPCHK(hipMemsetAsync(b->p, 0, sizeof(float) * N, c->stream));
    HIPCHK(hipMemsetAsync(c->flags.p, 0, sizeof(uint32_t) * NFLAG, c->stream));
    HIPCHK(hipMemsetAsync(c->bitmap.p, 0, sizeof(uint32_t) * c->nwords, c->stream));
    HIPCHK(hipMemsetAsync(c->ctl.p + C_BPC, 0, sizeof(uint32_t) * (C_CTL_WORDS - C_BPC), c->stream));
    // control block: counts zero, epoch advanced past every stored tag, word 15 = bitmap words
    uint32_t ctl[NCTL];
    HIPCHK(hipMemcpyAsync(ctl, c->ctl.p, sizeof(ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t ep = ctl[C_EP] + 4;
    std::memset(ctl, 0, sizeof(ctl));
    ctl[C_EP] = ep;
    ctl[15] = c->nwords;
    HIPCHK(hipMemcpyAsync(c->ctl.p, ctl, sizeof(ctl), hipMemcpyHostToDevice, c->stream));
    c->in = 0;
    std::fill(c->src.begin(), c->src.end(), 0.0);
    HIPCHK(hipStreamSynchronize(c->stream));
    return SM2D_OK;
}

template <class F>
int dispatch(sm2d_ctx* c, F&& f) {
    if (c->layout == L_REF) return f(LayoutRef{});
    return f(LayoutCfg{});
}

int check_ctx(sm2d_ctx* c) {
    if (!c) return err("null context");
    if (!c->tables) return err("tables not uploaded (sm2d_lut_upload)");
    return SM2D_OK;
}

int download_tallies(sm2d_ctx* c, double* t, int spot) {
    const size_t Ns = c->Ns, off = (size_t)spot * Ns;
    std::vector<float> f(Ns);
    HIPCHK(hipMemcpyAsync(t + 0 * Ns, c->edep.p + off, sizeof(double) * Ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(t + 2 * Ns, c->E_cut.p + off, sizeof(double) * Ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(t + 4 * Ns, c->E_nuc.p + off, sizeof(double) * Ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(t + 6 * Ns, c->E_bnd.p + off, sizeof(double) * Ns, hipMemcpyDeviceToHost, c->stream));
    const float* src[3] = {c->w_cut.p + off, c->w_nuc.p + off, c->w_bnd.p + off};
    const int dst[3] = {1, 3, 5};
    for (int k = 0; k < 3; ++k) {
        HIPCHK(hipMemcpyAsync(f.data(), src[k], sizeof(float) * Ns, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < Ns; ++i) t[dst[k] * Ns + i] = f[i];
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return SM2D_OK;
}

// Sum over spots (FP64, spot order) and optional FP32 copies of every spot's EdepC.
__global__ void k_edep_reduce(const double* edep, int n_spots, int Ns, double* sum, float* per) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < Ns; i += gridDim.x * blockDim.x) {
        double a = 0.0;
        for (int s = 0; s < n_spots; ++s) {
            double v = edep[(size_t)s * Ns + i];
            a += v;
            if (per) per[(size_t)s * Ns + i] = (float)v;
        }
        sum[i] = a;
    }
}

// Totals of the seven cumulative tallies over the first M cells (the spots of the last run), in
// FP64 and a fixed order: workgroup p of tally k sums its contiguous share of the cells with a
// fixed tree, then one workgroup adds the TSUM_PARTS partials of each tally in order.
constexpr int TSUM_PARTS = 512;

__device__ inline double tally_value(const Tally& t, int k, size_t i) {
    switch (k) {
        case 0: return t.edep[i];
        case 1: return (double)t.w_cut[i];
        case 2: return t.E_cut[i];
        case 3: return (double)t.w_nuc[i];
        case 4: return t.E_nuc[i];
        case 5: return (double)t.w_bnd[i];
        default: return t.E_bnd[i];
    }
}

__global__ __launch_bounds__(256) void k_tally_part(Tally t, size_t M, double* part) {
    __shared__ double sh[256];
    const int k = (int)blockIdx.y, p = (int)blockIdx.x;
    const size_t per = (M + TSUM_PARTS - 1) / TSUM_PARTS, i0 = (size_t)p * per, i1 = i0 + per < M ? i0 + per : M;
    double a = 0.0;
    for (size_t i = i0 + threadIdx.x; i < i1; i += 256) a += tally_value(t, k, i);
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)k * TSUM_PARTS + p] = sh[0];
}

__global__ __launch_bounds__(64) void k_tally_final(double* part) {
    const int k = (int)threadIdx.x;
    if (k >= 7) return;
    double a = 0.0;
    for (int p = 0; p < TSUM_PARTS; ++p) a += part[(size_t)k * TSUM_PARTS + p];
    part[(size_t)7 * TSUM_PARTS + k] = a;
}

// Capacities: listed cells per phase <= 48 rows of every spot (24 in a batch, whose spots are
// not all at their widest at once); occupied cells and destinations <= 5x that (the destination
// list in NSH hashed segments, each with room for twice its share and 64 more); payload slots
// <= 16 (8 in a batch) per listed cell, at least 64 Ki.  An overflow stops the transport and
// is reported as SM2D_ERR_CAPACITY.
void plan_caps(sm2d_ctx* c) {
    const size_t N = (size_t)c->N, S = (size_t)c->S;
    const size_t rows = S > 1 ? 24 : 48, per_listed = S > 1 ? 8 : 16;
    c->lcap = (uint32_t)std::min<size_t>(N, S * rows * c->g.Nx + 1024);
    c->ocap = (uint32_t)std::min<size_t>(N, 5ull * c->lcap);
    c->dcap = (uint32_t)((2 * ((size_t)c->ocap / NSH) + 64) * NSH);
    c->pcap = (uint32_t)std::min<size_t>(N * KB, std::max<size_t>(65536, per_listed * c->lcap));
    c->nwords = (uint32_t)((N + 31) / 32);
}

enum BufKind { BK_PSI, BK_BUCKET, BK_CELL, BK_OTHER, NBK };

// Every device buffer of a context with its element count and kind, in allocation order.
// `a(buffer, count, kind)` allocates it (sm2d_ctx_create) or only counts its bytes
// (sm2d_memory_estimate), so the estimate is the allocation itself.
template <class A>
int each_buffer(sm2d_ctx* c, A&& a) {
    const size_t N = (size_t)c->N, Ns = (size_t)c->Ns, S = (size_t)c->S;
    const size_t NT = (size_t)c->g.N_theta, NE = (size_t)c->g.N_E, Nz = (size_t)c->g.Nz;
    for (int k = 0; k < 2; ++k) {
        HIPCHK(a(c->bid[k], N * KB, BK_PSI));
        HIPCHK(a(c->ns[k], N, BK_PSI));
        HIPCHK(a(c->sref[k], N * KB, BK_PSI));
        HIPCHK(a(c->pcount[k], NSH * SH_STRIDE, BK_PSI));
        HIPCHK(a(c->pool[k], (size_t)c->pcap * c->LB, BK_PSI));
        HIPCHK(a(c->occ[k], c->ocap, BK_PSI));
    }
    HIPCHK(a(c->bitmap, c->nwords, BK_OTHER));
    HIPCHK(a(c->bbid, (size_t)c->lcap * 4 * KB, BK_BUCKET));
    HIPCHK(a(c->bn, (size_t)c->lcap * 4, BK_BUCKET));
    HIPCHK(a(c->bsref, (size_t)c->lcap * 4 * KB, BK_BUCKET));
    HIPCHK(a(c->bpool, (size_t)c->pcap * c->LB, BK_BUCKET));
    HIPCHK(a(c->bftag, c->lcap, BK_BUCKET));
    HIPCHK(a(c->bflist, c->lcap, BK_BUCKET));
    for (DBuf<uint32_t>* b : {&c->maskv, &c->ltag, &c->lpos, &c->dtag}) HIPCHK(a(*b, N, BK_CELL));
    HIPCHK(a(c->dlist, (size_t)c->dcap, BK_OTHER));
    HIPCHK(a(c->act, c->ocap, BK_OTHER));
    HIPCHK(a(c->crs, c->ocap, BK_OTHER));
    HIPCHK(a(c->ctl, C_CTL_WORDS, BK_OTHER));
    HIPCHK(a(c->gmoved, (size_t)2 * CHUNK * NSH * SH_STRIDE, BK_OTHER));
    HIPCHK(a(c->tcount, 2 * ((size_t)c->ocap / 256 + 2), BK_OTHER));
    HIPCHK(a(c->tcount2, (size_t)c->nwords / 256 + 2, BK_OTHER));
    HIPCHK(a(c->flags, NFLAG, BK_OTHER));
    HIPCHK(a(c->flag, c->ocap, BK_OTHER));
    for (DBuf<double>* b : {&c->edep, &c->E_cut, &c->E_nuc, &c->E_bnd, &c->k4w, &c->k4e})
        HIPCHK(a(*b, N, BK_CELL));
    HIPCHK(a(c->win, c->ocap, BK_OTHER));
    HIPCHK(a(c->ein, c->ocap, BK_OTHER));
    HIPCHK(a(c->acc_c, (size_t)c->ocap * NACC, BK_OTHER));
    HIPCHK(a(c->acc_f, (size_t)c->ocap * NACC, BK_OTHER));
    HIPCHK(a(c->pe, (size_t)c->ocap * NR1, BK_OTHER));
    HIPCHK(a(c->part, ((size_t)c->ocap / 256 + 2) * NR1, BK_OTHER));
    HIPCHK(a(c->k5v, NV, BK_OTHER));
    HIPCHK(a(c->hist, (size_t)HROWS * NV, BK_OTHER));
    HIPCHK(a(c->edep_sum, Ns, BK_OTHER));
    HIPCHK(a(c->tsum, 7 * (size_t)(TSUM_PARTS + 1), BK_OTHER));
    HIPCHK(a(c->accpart, 2 * (size_t)acc_tcap(c) * NACCP, BK_OTHER));
    for (DBuf<float>* b : {&c->w_cut, &c->w_nuc, &c->w_bnd}) HIPCHK(a(*b, N, BK_CELL));
    HIPCHK(a(c->spots, S, BK_OTHER));
    HIPCHK(a(c->injacc, S * 6, BK_OTHER));
    HIPCHK(a(c->te, NT + 1, BK_OTHER));
    HIPCHK(a(c->ee, NE + 1, BK_OTHER));
    for (DBuf<float>* b : {&c->R, &c->Sp, &c->lE, &c->lR, &c->lS}) HIPCHK(a(*b, NE, BK_OTHER));
    HIPCHK(a(c->feA, S * Nz, BK_OTHER));
    HIPCHK(a(c->feB, S * Nz, BK_OTHER));
    HIPCHK(a(c->sx0, S, BK_OTHER));
    HIPCHK(a(c->rho, Nz, BK_OTHER));
    HIPCHK(a(c->x0r, Nz, BK_OTHER));
    return SM2D_OK;
}

// Grid and layout checks of a context (shared by the estimate and the creation).
int check_grid(const sm2d_grid* g, int n_spots, LayoutId* lid) {
    if (!g || g->struct_size != sizeof(sm2d_grid)) return err("sm2d_ctx_create: grid struct_size");
    if (g->Nx <= 0 || g->Nz <= 0 || g->N_theta <= 0 || g->N_E <= 0 || n_spots <= 0 ||
        g->N_theta > 4095 * g->n_theta_local)
        return err("sm2d_ctx_create: bad grid");
    // the transport stages the energy tables in LDS next to its payload rows (64 KiB per wave)
    if (g->N_E > 1100) return err("sm2d_ctx_create: N_E above 1100 is not supported", SM2D_ERR_UNSUPPORTED);
    if (g->n_theta_local == 4 && g->n_e_local == 2 && g->n_x_sub == 8 && g->n_z_sub == 4) *lid = L_REF;
    else if (g->n_theta_local == 8 && g->n_e_local == 4 && g->n_x_sub == 4 && g->n_z_sub == 1) *lid = L_CFG;
    else return err("sm2d_ctx_create: local-bin layout must be (4,2,8,4) or (8,4,4,1)");
    if ((size_t)g->Nx * g->Nz * (size_t)n_spots >= (1ull << 31)) return err("sm2d_ctx_create: too many cells");
    return SM2D_OK;
}

int create(int device, const sm2d_grid* g, int n_spots, sm2d_ctx** out) {
    if (!out) return err("sm2d_ctx_create: out is NULL");
    LayoutId lid;
    int rc_alloc = check_grid(g, n_spots, &lid);
    if (rc_alloc) return rc_alloc;
    const size_t Ns = (size_t)g->Nx * g->Nz, N = Ns * (size_t)n_spots;
    HIPCHK(hipSetDevice(device));
    std::unique_ptr<sm2d_ctx> c(new sm2d_ctx());
    c->device = device;
    c->g = *g;
    c->layout = lid;
    c->LB = g->n_theta_local * g->n_e_local * g->n_x_sub * g->n_z_sub;
    c->S = n_spots;
    c->Ns = (int)Ns;
    c->N = (int)N;
    plan_caps(c.get());
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {   // the 4-wave transport's dynamic LDS (energy tables) comes on top of its static LDS
        const int dyn = (int)transport_dyn_lds(g->N_E, 4);
        const void* fns[4] = {(const void*)&k_transport<LayoutRef, true, 4>,
                              (const void*)&k_transport<LayoutRef, false, 4>,
                              (const void*)&k_transport<LayoutCfg, true, 4>,
                              (const void*)&k_transport<LayoutCfg, false, 4>};
        for (int k = (lid == L_REF ? 0 : 2); k < (lid == L_REF ? 2 : 4); ++k)
            HIPCHK(hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
    }
    c->bytes = 0;
    rc_alloc = each_buffer(c.get(), [&](auto& b, size_t n, int) {
        c->bytes += b.bytes_for(n);
        return b.alloc(n);
    });
    if (rc_alloc) return rc_alloc;
    c->src.assign((size_t)n_spots * 7, 0.0);
    HIPCHK(hipMemsetAsync(c->maskv.p, 0, sizeof(uint32_t) * N, c->stream));
    HIPCHK(hipMemsetAsync(c->bftag.p, 0, sizeof(uint32_t) * c->lcap, c->stream));
    HIPCHK(hipMemsetAsync(c->ltag.p, 0, sizeof(uint32_t) * N, c->stream));
    HIPCHK(hipMemsetAsync(c->ctl.p, 0, sizeof(uint32_t) * C_CTL_WORDS, c->stream));
    HIPCHK(hipMemsetAsync(c->dtag.p, 0, sizeof(uint32_t) * N, c->stream));
    HIPCHK(hipMemsetAsync(c->feA.p, 0, sizeof(float) * n_spots * g->Nz, c->stream));
    HIPCHK(hipMemsetAsync(c->feB.p, 0, sizeof(float) * n_spots * g->Nz, c->stream));
    HIPCHK(hipMemsetAsync(c->sx0.p, 0, sizeof(float) * n_spots, c->stream));
    std::vector<float> ones(g->Nz, 1.0f);
    HIPCHK(hipMemcpyAsync(c->rho.p, ones.data(), sizeof(float) * g->Nz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->x0r.p, ones.data(), sizeof(float) * g->Nz, hipMemcpyHostToDevice, c->stream));
    c->hint = (uint32_t)std::min<size_t>(c->lcap, (size_t)n_spots * g->Nx);
    c->hint_ph[0] = c->hint_ph[1] = c->hint;
    c->ev.resize((size_t)CHUNK * sm2d_ctx::EPI);
    for (hipEvent_t& e : c->ev) HIPCHK(hipEventCreate(&e));
    if (const char* f = std::getenv("SM2D_TRANSPORT_WAVES")) c->tune.forced = std::atoi(f);
    if (const char* f = std::getenv("SM2D_GRAPHS")) c->graphs_on = std::atoi(f) != 0;
    int rc = zero_state(c.get());
    if (rc) return rc;
    *out = c.release();
    return SM2D_OK;
}

// FE moments + beam width of spot slot s from the host precompute (init_pipeline_state,
// k1k6_pipeline.cu:1061-1212).
int upload_fe_for(sm2d_ctx* c, int s, const sm2d_spot& sp, std::vector<float>* sigma_total) {
    const int Nz = c->g.Nz;
    std::vector<float> A(Nz), B(Nz), st(Nz);
    int rc = sm2d_fe_precompute(sp.E0, sp.sigma_x, Nz, c->g.dz, A.data(), B.data(), st.data());
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->feA.p + (size_t)s * Nz, A.data(), sizeof(float) * Nz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->feB.p + (size_t)s * Nz, B.data(), sizeof(float) * Nz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->sx0.p + s, &sp.sigma_x, sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (sigma_total) *sigma_total = st;
    return SM2D_OK;
}

// The moments of n spots (host precompute, spread over host threads), then one upload.
int upload_fe_batch(sm2d_ctx* c, const sm2d_spot* sp, int n) {
    const int Nz = c->g.Nz;
    std::vector<float> A((size_t)n * Nz), B((size_t)n * Nz), sx(n);
    std::vector<int> rcs(n, SM2D_OK);
    const int nt = std::max(1, std::min<int>({n, 16, (int)std::thread::hardware_concurrency()}));
    auto work = [&](int t) {
        std::vector<float> st(Nz);
        for (int s = t; s < n; s += nt)
            rcs[s] = sm2d_fe_precompute(sp[s].E0, sp[s].sigma_x, Nz, c->g.dz, A.data() + (size_t)s * Nz,
                                        B.data() + (size_t)s * Nz, st.data());
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) th.emplace_back(work, t);
        for (std::thread& x : th) x.join();
    }
    for (int s = 0; s < n; ++s) {
        if (rcs[s]) {  // again on this thread, which keeps the error message
            std::vector<float> st(Nz);
            return sm2d_fe_precompute(sp[s].E0, sp[s].sigma_x, Nz, c->g.dz, A.data(), B.data(), st.data());
        }
        sx[s] = sp[s].sigma_x;
    }
    HIPCHK(hipMemcpyAsync(c->feA.p, A.data(), sizeof(float) * A.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->feB.p, B.data(), sizeof(float) * B.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->sx0.p, sx.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SM2D_OK;
}

}  // namespace

#include "sm2d_surrogate.hip"
#include "sm2d_noalloc.hip"

extern "C" {

const char* sm2d_last_error(void) { return sm2d_host::error(); }
int sm2d_abi_version(void) { return SM2D_ABI_VERSION; }

int sm2d_ctx_create(int device, const sm2d_grid* g, sm2d_ctx** out) { return create(device, g, 1, out); }

int sm2d_ctx_create_batch(int device, const sm2d_grid* g, int n_spots, sm2d_ctx** out) {
    return create(device, g, n_spots, out);
}

int sm2d_stage_clocks(int device, double* out32, int reset) {
#ifdef SM2D_STAGE_CLOCKS
    if (!out32) return err("sm2d_stage_clocks: no output");
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipDeviceSynchronize());
    unsigned long long v[32];
    HIPCHK(hipMemcpyFromSymbol(v, HIP_SYMBOL(stage_clk), sizeof(v)));
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
    for (int k = 0; k < 32; ++k) out32[k] = k == 15 ? (double)v[k] : khz > 0 ? (double)v[k] * 1e3 / (double)khz : 0.0;
    if (reset) {
        std::memset(v, 0, sizeof(v));
        HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(stage_clk), v, sizeof(v)));
    }
    return SM2D_OK;
#else
    (void)device;
    (void)out32;
    (void)reset;
    return err("stage clocks need the diagnostic build (python -m sm_2d_amd.build --stage-clocks)",
               SM2D_ERR_UNSUPPORTED);
#endif
}

int sm2d_device_name(int device, char* buf, int capacity) {
    if (!buf || capacity <= 0) return err("sm2d_device_name: no buffer");
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    std::snprintf(buf, (size_t)capacity, "%s (%s)", p.name, p.gcnArchName);
    return SM2D_OK;
}

int sm2d_memory_estimate(const sm2d_grid* g, int n_spots, sm2d_memory_plan* out) {
    if (!out || out->struct_size != sizeof(sm2d_memory_plan)) return err("sm2d_memory_estimate: struct_size");
    LayoutId lid;
    int rc = check_grid(g, n_spots, &lid);
    if (rc) return rc;
    sm2d_ctx c;  // host-side planning copy: capacities only, nothing is allocated
    c.g = *g;
    c.LB = g->n_theta_local * g->n_e_local * g->n_x_sub * g->n_z_sub;
    c.S = n_spots;
    c.Ns = g->Nx * g->Nz;
    c.N = c.Ns * n_spots;
    plan_caps(&c);
    uint64_t kind[NBK] = {0, 0, 0, 0};
    rc = each_buffer(&c, [&](auto& b, size_t n, int k) {
        kind[k] += b.bytes_for(n);
        return hipSuccess;
    });
    if (rc) return rc;
    const uint32_t keep = out->struct_size;
    *out = sm2d_memory_plan{};
    out->struct_size = keep;
    out->listed_cells = c.lcap;
    out->occupied_cells = c.ocap;
    out->payload_slots = c.pcap;
    out->psi_bytes = kind[BK_PSI];
    out->bucket_bytes = kind[BK_BUCKET];
    out->cell_bytes = kind[BK_CELL];
    out->other_bytes = kind[BK_OTHER];
    out->total_bytes = kind[0] + kind[1] + kind[2] + kind[3];
    // the reference's dense store for the same grid (estimate_dense_k1k6_memory,
    // src/perf/memory_preflight.cpp:22-105) with this layout's local bins: two full ψ
    // buffers, four buckets per cell, 114 B of pipeline state per cell (2 masks, 2 lists and
    // the 14 current/previous tally arrays and 2 weight sums of memory_preflight.cpp:22-39),
    // two counters, the audit report, the edges, the LUT and 64 MiB of runtime overhead
    const uint64_t cells = (uint64_t)c.N, LB = (uint64_t)c.LB;
    const uint64_t psic_cell = 32ull * (4 + LB * 4), face = 32ull * (4 + 2 + LB * 4) + 12;
    out->dense_bytes = cells * (2 * psic_cell + 4 * face + 114) + 8 + 256 +
                       4ull * (g->N_theta + 1) + 4ull * (g->N_E + 1) + 4ull * (6ull * g->N_E + 1) +
                       (64ull << 20);
    out->ok = 1;
    return SM2D_OK;
}

int sm2d_memory_preflight(int device, const sm2d_grid* g, int n_spots, float margin,
                          sm2d_memory_plan* out) {
    int rc = sm2d_memory_estimate(g, n_spots, out);
    if (rc) return rc;
    if (!(margin > 0.0f && margin <= 1.0f)) return err("sm2d_memory_preflight: margin must be in (0,1]");
    HIPCHK(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    out->free_bytes = fr;
    out->usable_bytes = (uint64_t)std::floor((double)fr * (double)margin);
    out->ok = out->total_bytes <= out->usable_bytes;
    return SM2D_OK;
}

int sm2d_ctx_memory(sm2d_ctx* c, uint64_t* bytes) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!bytes) return err("sm2d_ctx_memory: bytes is NULL");
    *bytes = c->bytes;
    return SM2D_OK;
}

int sm2d_ctx_destroy(sm2d_ctx* c) {
    if (!c) return SM2D_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    c->graphs_clear();
    hipStream_t s = c->stream;
    delete c;
    if (s) (void)hipStreamDestroy(s);
    return SM2D_OK;
}

int sm2d_lut_upload(sm2d_ctx* c, const float* te, const float* ee, const float* R, const float* S,
                    const float* lE, const float* lR, const float* lS) {
    if (!c || !te || !ee || !R || !S || !lE || !lR || !lS) return err("sm2d_lut_upload: null argument");
    const int NT = c->g.N_theta, NE = c->g.N_E;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->te.p, te, sizeof(float) * (NT + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ee.p, ee, sizeof(float) * (NE + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->R.p, R, sizeof(float) * NE, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->Sp.p, S, sizeof(float) * NE, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->lE.p, lE, sizeof(float) * NE, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->lR.p, lR, sizeof(float) * NE, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->lS.p, lS, sizeof(float) * NE, hipMemcpyHostToDevice, c->stream));
    c->h_ee.assign(ee, ee + NE + 1);
    c->h_R.assign(R, R + NE);
    c->h_lR.assign(lR, lR + NE);
    c->h_lE.assign(lE, lE + NE);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->tables = true;
    return SM2D_OK;
}

int sm2d_fe_upload(sm2d_ctx* c, const float* A, const float* B) {
    if (!c) return err("null context");
    const int Nz = c->g.Nz;
    HIPCHK(hipSetDevice(c->device));
    if (!A || !B) {
        HIPCHK(hipMemsetAsync(c->feA.p, 0, sizeof(float) * Nz, c->stream));
        HIPCHK(hipMemsetAsync(c->feB.p, 0, sizeof(float) * Nz, c->stream));
    } else {
        HIPCHK(hipMemcpyAsync(c->feA.p, A, sizeof(float) * Nz, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->feB.p, B, sizeof(float) * Nz, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return SM2D_OK;
}

int sm2d_material_upload(sm2d_ctx* c, const float* rho, const float* x0r) {
    if (!c) return err("null context");
    const int Nz = c->g.Nz;
    std::vector<float> ones(Nz, 1.0f);
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->rho.p, rho ? rho : ones.data(), sizeof(float) * Nz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->x0r.p, x0r ? x0r : ones.data(), sizeof(float) * Nz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SM2D_OK;
}

int sm2d_reset(sm2d_ctx* c) {
    int rc = check_ctx(c);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    return zero_state(c);
}

int sm2d_inject(sm2d_ctx* c, const sm2d_spot* sp, sm2d_audit* au) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!sp || sp->struct_size != sizeof(sm2d_spot)) return err("spot struct_size");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->sx0.p, &sp->sigma_x, sizeof(float), hipMemcpyHostToDevice, c->stream));
    rc = dispatch(c, [&](auto ly) { return inject_all<decltype(ly)>(c, sp, 1); });
    if (rc) return rc;
    if (au) {
        au->W_source_injected = c->src[0];
        au->W_source_out_of_grid = c->src[1];
        au->W_source_slot_drop = c->src[2];
        au->E_source_injected = c->src[3];
        au->E_source_out_of_grid = c->src[4];
        au->E_source_slot_drop = c->src[5];
        au->E_representation_loss = c->src[6];
    }
    return SM2D_OK;
}

int sm2d_psi_load(sm2d_ctx* c, int n, const int32_t* cell, const int32_t* slot, const uint32_t* bid,
                  const int32_t* lidx, const float* w) {
    int rc = check_ctx(c);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    const size_t N = c->N;
    const int LB = c->LB;
    std::vector<uint32_t> hb(N * KB, SM2D_EMPTY), hn(N, 0), hs(N * KB, 0);
    std::vector<float> pool;
    std::vector<int64_t> where(N * KB, -1);
    for (int i = 0; i < n; ++i) {
        int cl = cell[i], s = slot[i];
        if (cl < 0 || (size_t)cl >= N || s < 0 || s >= KB || lidx[i] < 0 || lidx[i] >= LB)
            return err("psi_load: row out of range");
        // the bin must lie on the grid: a partial last block has local bins past N_theta or
        // N_E, which no emission produces and whose energy centre has no upper edge
        const int TL = c->g.n_theta_local, EL = c->g.n_e_local;
        if ((bid[i] >> 24) != 0 || (int)(bid[i] & 0xFFF) * TL + lidx[i] % TL >= c->g.N_theta ||
            (int)((bid[i] >> 12) & 0xFFF) * EL + (lidx[i] / TL) % EL >= c->g.N_E)
            return err("psi_load: bin outside the grid");
        size_t k = (size_t)cl * KB + s;
        if (hb[k] != SM2D_EMPTY && hb[k] != bid[i]) return err("psi_load: rows of one slot carry different bids");
        if (where[k] < 0) {
            where[k] = (int64_t)(pool.size() / LB);
            pool.resize(pool.size() + LB, 0.0f);
        }
        hb[k] = bid[i];
        hn[cl] = std::max<uint32_t>(hn[cl], (uint32_t)s + 1);
        pool[(size_t)where[k] * LB + lidx[i]] = w[i];
    }
    std::vector<uint32_t> occ;
    for (size_t cl = 0; cl < N; ++cl) {
        for (uint32_t s = 0; s < hn[cl]; ++s) {
            size_t k = cl * KB + s;
            if (where[k] < 0) return err("psi_load: slots of a cell must be a prefix");
            hs[k] = (uint32_t)where[k];
        }
        if (hn[cl]) occ.push_back((uint32_t)cl);
    }
    // the loaded rows fill the pool's shards in order (row r lies in shard r / span)
    const uint32_t used = (uint32_t)(pool.size() / LB), span = c->pcap / NSH;
    if (used > span * NSH) return err("psi_load: pool capacity", SM2D_ERR_CAPACITY);
    std::vector<uint32_t> shard(NSH * SH_STRIDE, 0);
    for (int k = 0; k < NSH; ++k)
        shard[(size_t)k * SH_STRIDE] = std::min(span, used - std::min(used, (uint32_t)k * span));
    if (occ.size() > c->ocap) return err("psi_load: occupied-cell capacity", SM2D_ERR_CAPACITY);
    rc = zero_state(c);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->bid[0].p, hb.data(), sizeof(uint32_t) * hb.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ns[0].p, hn.data(), sizeof(uint32_t) * hn.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->sref[0].p, hs.data(), sizeof(uint32_t) * hs.size(), hipMemcpyHostToDevice, c->stream));
    if (!pool.empty())
        HIPCHK(hipMemcpyAsync(c->pool[0].p, pool.data(), sizeof(float) * pool.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->pcount[0].p, shard.data(), sizeof(uint32_t) * shard.size(), hipMemcpyHostToDevice,
                          c->stream));
    if (!occ.empty())
        HIPCHK(hipMemcpyAsync(c->occ[0].p, occ.data(), sizeof(uint32_t) * occ.size(), hipMemcpyHostToDevice, c->stream));
    uint32_t no = (uint32_t)occ.size();
    HIPCHK(hipMemcpyAsync(c->nocc(0), &no, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    seed_hints(c, no);
    HIPCHK(hipStreamSynchronize(c->stream));
    return SM2D_OK;
}

int sm2d_psi_dump(sm2d_ctx* c, uint32_t* bids, int cap, int32_t* cell, int32_t* slot, uint32_t* bid,
                  int32_t* lidx, float* w, int* nrows) {
    int rc = check_ctx(c);
    if (rc) return rc;
    HIPCHK(hipSetDevice(c->device));
    const size_t N = c->N;
    const int LB = c->LB, k = c->in;
    std::vector<uint32_t> hb(N * KB), hn(N), hs(N * KB), cnt(NSH * SH_STRIDE);
    HIPCHK(hipMemcpyAsync(hb.data(), c->bid[k].p, sizeof(uint32_t) * hb.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hn.data(), c->ns[k].p, sizeof(uint32_t) * hn.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hs.data(), c->sref[k].p, sizeof(uint32_t) * hs.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(cnt.data(), c->pcount[k].p, sizeof(uint32_t) * cnt.size(), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // the used rows of every shard, packed: row r of shard r / span sits at off[r / span] + r % span
    const uint32_t span = c->pcap / NSH;
    std::vector<size_t> off(NSH + 1, 0);
    for (int s = 0; s < NSH; ++s) off[s + 1] = off[s] + std::min(cnt[(size_t)s * SH_STRIDE], span);
    std::vector<float> pool(off[NSH] * LB);
    for (int s = 0; s < NSH; ++s)
        if (off[s + 1] > off[s])
            HIPCHK(hipMemcpyAsync(pool.data() + off[s] * LB, c->pool[k].p + (size_t)s * span * LB,
                                  sizeof(float) * (off[s + 1] - off[s]) * LB, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (bids) std::memcpy(bids, hb.data(), sizeof(uint32_t) * hb.size());
    int m = 0;
    for (size_t cl = 0; cl < N; ++cl)
        for (uint32_t s = 0; s < hn[cl]; ++s) {
            const uint32_t r = hs[cl * KB + s];
            if (r / span >= (uint32_t)NSH || r % span >= off[r / span + 1] - off[r / span])
                return err("psi_dump: slot payload outside the used pool");
            const float* v = pool.data() + (off[r / span] + r % span) * LB;
            for (int l = 0; l < LB; ++l) {
                if (v[l] == 0.0f) continue;
                if (m < cap) {
                    cell[m] = (int32_t)cl;
                    slot[m] = (int32_t)s;
                    bid[m] = hb[cl * KB + s];
                    lidx[m] = l;
                    w[m] = v[l];
                }
                ++m;
            }
        }
    *nrows = m;
    return SM2D_OK;
}

int sm2d_k1k2_mask_compact(sm2d_ctx* c, const sm2d_transport* tp, const uint8_t* mask_prev, uint8_t* mask_out,
                           int32_t* active, int* n_active, int32_t* coarse, int* n_coarse) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!tp || tp->struct_size != sizeof(sm2d_transport)) return err("transport struct_size");
    HIPCHK(hipSetDevice(c->device));
    uint32_t ctl[NCTL];
    HIPCHK(hipMemcpyAsync(ctl, c->ctl.p, sizeof(ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t ep = ctl[C_EP] + 1;  // the epoch k_begin is about to open
    std::vector<uint32_t> mv(c->N, 0);
    if (mask_prev)
        for (int i = 0; i < c->N; ++i) mv[i] = mask_prev[i] ? (((ep - 1u) << 2) | 1u) : 0u;
    HIPCHK(hipMemcpyAsync(c->maskv.p, mv.data(), sizeof(uint32_t) * c->N, hipMemcpyHostToDevice, c->stream));
    Tp P = make_tp(c, tp);
    k_ctl_reset<<<1, 1, 0, c->stream>>>(c->ctl.p);
    k_begin<<<1, 64, 0, c->stream>>>(c->ctl.p);
    const int in = c->in;
    dispatch(c, [&](auto ly) {
        launch_mask_compact<decltype(ly)>(c, P, in, (int)std::max<uint32_t>(c->hint_occ, 1024));
        return 0;
    });
    k_fill_u32<<<blocks_for(c->nwords, 256, 4096), 256, 0, c->stream>>>(c->bitmap.p, 0u, c->nwords);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctl, c->ctl.p, sizeof(ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t na = ctl[C_NACT], nc = ctl[C_NCRS];
    std::vector<uint32_t> a(na ? na : 1), k(nc ? nc : 1);
    HIPCHK(hipMemcpyAsync(mv.data(), c->maskv.p, sizeof(uint32_t) * c->N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(a.data(), c->act.p, sizeof(uint32_t) * na, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(k.data(), c->crs.p, sizeof(uint32_t) * nc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (mask_out)
        for (int i = 0; i < c->N; ++i) mask_out[i] = ((mv[i] >> 2) == ep && (mv[i] & 1u)) ? 1 : 0;
    for (uint32_t i = 0; i < na; ++i) active[i] = (int32_t)a[i];
    for (uint32_t i = 0; i < nc; ++i) coarse[i] = (int32_t)k[i];
    *n_active = (int)na;
    *n_coarse = (int)nc;
    return SM2D_OK;
}

int sm2d_run_iterations(sm2d_ctx* c, const sm2d_transport* tp, int n_iter, int src_terms, double* tallies,
                        sm2d_k5_report* last, sm2d_audit* au) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!tp || tp->struct_size != sizeof(sm2d_transport)) return err("transport struct_size");
    HIPCHK(hipSetDevice(c->device));
    if (au) {
        std::memset(au, 0, sizeof(*au));
        au->struct_size = sizeof(*au);
    }
    HIPCHK(hipMemcpyAsync(c->sx0.p, &tp->sigma_x_initial, sizeof(float), hipMemcpyHostToDevice, c->stream));
    rc = dispatch(c, [&](auto ly) { return run_iters<decltype(ly)>(c, tp, n_iter, src_terms, last, au); });
    if (rc) return rc;
    if (au) {
        uint32_t f[NFLAG];
        HIPCHK(hipMemcpy(f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost));
        au->cross_cell_emissions = f[F_CROSS];
    }
    if (tallies) return download_tallies(c, tallies, 0);
    return SM2D_OK;
}

int sm2d_k5_audit(sm2d_ctx* c, const double* td, double Wso, double Wss, double Eso, double Ess, double Wd,
                  double Ed, int src, sm2d_k5_report* out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!td || !out) return err("sm2d_k5_audit: null argument");
    HIPCHK(hipSetDevice(c->device));
    double Wi, Ei, Wo, Eo;
    rc = dispatch(c, [&](auto ly) { return buffer_totals<decltype(ly)>(c, c->in, Wi, Ei); });
    if (rc) return rc;
    rc = dispatch(c, [&](auto ly) { return buffer_totals<decltype(ly)>(c, 1 - c->in, Wo, Eo); });
    if (rc) return rc;
    const size_t N = c->Ns;
    double v[NV] = {};
    double* a = v + NR1;
    for (size_t i = 0; i < N; ++i) {
        a[A_EDEP] += std::max(0.0, td[0 * N + i]);
        a[A_WCUT] += std::max(0.0, td[1 * N + i]);
        a[A_ECUT] += std::max(0.0, td[2 * N + i]);
        a[A_WNUC] += std::max(0.0, td[3 * N + i]);
        a[A_ENUC] += std::max(0.0, td[4 * N + i]);
        a[A_WBND] += std::max(0.0, td[5 * N + i]);
        a[A_EBND] += std::max(0.0, td[6 * N + i]);
    }
    v[R_WIN] = Wi;
    v[R_EIN] = Ei;
    v[R_WOUT] = Wo;
    v[R_EOUT] = Eo;
    a[A_DROPW] = Wd;
    a[A_DROPE] = Ed;
    fill_report(v, Wso, Wss, Eso, Ess, src, out);
    return SM2D_OK;
}

int sm2d_swap(sm2d_ctx* c) {
    if (!c) return err("null context");
    c->in = 1 - c->in;
    return SM2D_OK;
}

static int finish_audit(sm2d_ctx* c, sm2d_audit* au, const sm2d_audit& it, int n_spots) {
    au->iterations = it.iterations;
    au->terminated = it.terminated;
    au->W_fail_iters = it.W_fail_iters;
    au->E_fail_iters = it.E_fail_iters;
    au->max_W_error = it.max_W_error;
    au->max_E_error = it.max_E_error;
    au->W_transport_drop = it.W_transport_drop;
    au->E_transport_drop = it.E_transport_drop;
    au->E_audit_residual = it.E_audit_residual;
    au->max_cell_W_closure = it.max_cell_W_closure;
    au->max_cell_E_closure = it.max_cell_E_closure;
    au->cell_bin_updates = it.cell_bin_updates;
    au->processed_components = it.processed_components;
    au->processed_cells = it.processed_cells;
    au->surrogate_active = it.surrogate_active;
    au->surrogate_fallback = it.surrogate_fallback;
    au->fine_cell_bin_updates = it.fine_cell_bin_updates;
    au->surrogate_substeps = it.surrogate_substeps;
    for (int s = 0; s < n_spots; ++s) {
        au->W_source_injected += c->src[s * 7 + 0];
        au->W_source_out_of_grid += c->src[s * 7 + 1];
        au->W_source_slot_drop += c->src[s * 7 + 2];
        au->E_source_injected += c->src[s * 7 + 3];
        au->E_source_out_of_grid += c->src[s * 7 + 4];
        au->E_source_slot_drop += c->src[s * 7 + 5];
        au->E_representation_loss += c->src[s * 7 + 6];
    }
    if (!au->terminated) {
        int rc = dispatch(c, [&](auto ly) { return buffer_totals<decltype(ly)>(c, c->in, au->W_leftover, au->E_leftover); });
        if (rc) return rc;
    }
    uint32_t f[NFLAG];
    HIPCHK(hipMemcpy(f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost));
    au->cross_cell_emissions = f[F_CROSS];
    return SM2D_OK;
}

// Totals of the per-cell tallies of spots [0, n) (FP64, on the host from a device reduction).
static int tally_totals(sm2d_ctx* c, int n, double* s7) {
    k_tally_part<<<dim3(TSUM_PARTS, 7), 256, 0, c->stream>>>(c->tally(), (size_t)n * c->Ns, c->tsum.p);
    k_tally_final<<<1, 64, 0, c->stream>>>(c->tsum.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(s7, c->tsum.p + (size_t)7 * TSUM_PARTS, sizeof(double) * 7, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SM2D_OK;
}

int sm2d_run_spot(sm2d_ctx* c, const sm2d_spot* sp, const sm2d_transport* tp, double* edep_out, sm2d_audit* au) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!tp || tp->struct_size != sizeof(sm2d_transport)) return err("transport struct_size");
    if (!sp || sp->struct_size != sizeof(sm2d_spot)) return err("spot struct_size");
    if (tp->max_iterations <= 0) return err("max_iterations must be positive");
    sm2d_audit local;
    if (!au) au = &local;
    std::memset(au, 0, sizeof(*au));
    au->struct_size = sizeof(*au);
    HIPCHK(hipSetDevice(c->device));
    rc = zero_state(c);
    if (rc) return rc;
    rc = upload_fe_for(c, 0, *sp, nullptr);
    if (rc) return rc;
    rc = dispatch(c, [&](auto ly) { return inject_all<decltype(ly)>(c, sp, 1); });
    if (rc) return rc;
    sm2d_audit it;
    std::memset(&it, 0, sizeof(it));
    rc = dispatch(c, [&](auto ly) { return run_iters<decltype(ly)>(c, tp, tp->max_iterations, 1, nullptr, &it); });
    if (rc) return rc;
    double s7[7];
    rc = tally_totals(c, 1, s7);
    if (rc) return rc;
    au->E_dep = s7[0]; au->W_cutoff = s7[1]; au->E_cutoff = s7[2]; au->W_nuclear = s7[3];
    au->E_nuclear = s7[4]; au->W_boundary = s7[5]; au->E_boundary = s7[6];
    if (edep_out)
        HIPCHK(hipMemcpy(edep_out, c->edep.p, sizeof(double) * c->Ns, hipMemcpyDeviceToHost));
    return finish_audit(c, au, it, 1);
}

int sm2d_run_batch(sm2d_ctx* c, const sm2d_spot* spots, int n, const sm2d_transport* tp, double* edep_sum,
                   float* edep_per_spot, sm2d_audit* au) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!tp || tp->struct_size != sizeof(sm2d_transport)) return err("transport struct_size");
    if (!spots || n <= 0 || n > c->S) return err("sm2d_run_batch: 1 <= n_spots <= the context's spot slots");
    for (int s = 0; s < n; ++s)
        if (spots[s].struct_size != sizeof(sm2d_spot)) return err("spot struct_size");
    if (tp->max_iterations <= 0) return err("max_iterations must be positive");
    sm2d_audit local;
    if (!au) au = &local;
    std::memset(au, 0, sizeof(*au));
    au->struct_size = sizeof(*au);
    HIPCHK(hipSetDevice(c->device));
    rc = zero_state(c);
    if (rc) return rc;
    rc = upload_fe_batch(c, spots, n);
    if (rc) return rc;
    rc = dispatch(c, [&](auto ly) { return inject_all<decltype(ly)>(c, spots, n); });
    if (rc) return rc;
    sm2d_audit it;
    std::memset(&it, 0, sizeof(it));
    rc = dispatch(c, [&](auto ly) { return run_iters<decltype(ly)>(c, tp, tp->max_iterations, 1, nullptr, &it); });
    if (rc) return rc;
    DBuf<float> per;
    if (edep_per_spot) HIPCHK(per.alloc((size_t)n * c->Ns));
    k_edep_reduce<<<blocks_for(c->Ns, 256, 4096), 256, 0, c->stream>>>(c->edep.p, n, c->Ns, c->edep_sum.p,
                                                                       edep_per_spot ? per.p : nullptr);
    HIPCHK(hipGetLastError());
    if (edep_sum)
        HIPCHK(hipMemcpyAsync(edep_sum, c->edep_sum.p, sizeof(double) * c->Ns, hipMemcpyDeviceToHost, c->stream));
    if (edep_per_spot)
        HIPCHK(hipMemcpyAsync(edep_per_spot, per.p, sizeof(float) * n * c->Ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->batch_n = n;
    double s7[7];
    rc = tally_totals(c, n, s7);
    if (rc) return rc;
    au->E_dep = s7[0]; au->W_cutoff = s7[1]; au->E_cutoff = s7[2]; au->W_nuclear = s7[3];
    au->E_nuclear = s7[4]; au->W_boundary = s7[5]; au->E_boundary = s7[6];
    return finish_audit(c, au, it, n);
}

int sm2d_batch_spot_edep(sm2d_ctx* c, int s, double* edep_out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!edep_out || s < 0 || s >= c->batch_n) return err("sm2d_batch_spot_edep: spot index outside the last batch");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(edep_out, c->edep.p + (size_t)s * c->Ns, sizeof(double) * c->Ns, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SM2D_OK;
}

int sm2d_set_profiling(sm2d_ctx* c, int on) {
    if (!c) return err("null context");
    HIPCHK(hipSetDevice(c->device));
    c->prof = on != 0;
    std::memset(&c->timing, 0, sizeof(c->timing));
    c->timing.struct_size = sizeof(c->timing);
    c->timing.enabled = c->prof ? 1 : 0;
    return SM2D_OK;
}

int sm2d_set_transport_waves(sm2d_ctx* c, int waves) {
    if (!c) return err("null context");
    if (waves != 0 && waves != 1 && waves != 2 && waves != 4) return err("waves per cell must be 0, 1, 2 or 4");
    c->tune.forced = waves;
    return SM2D_OK;
}

int sm2d_set_gather_rows(sm2d_ctx* c, int kib) {
    if (!c) return err("null context");
    if (kib != 0 && kib != 4 && kib != 8 && kib != 16) return err("gather rows must be 0, 4, 8 or 16 KiB");
    c->tune.forced_kib = kib;
    return SM2D_OK;
}

int sm2d_set_graphs(sm2d_ctx* c, int on) {
    if (!c) return err("null context");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->graphs_on = on != 0;
    if (!c->graphs_on) c->graphs_clear();
    return SM2D_OK;
}

int sm2d_tuner_report(sm2d_ctx* c, char* buf, int cap) {
    if (!c || !buf || cap <= 0) return err("sm2d_tuner_report: null argument");
    std::string o = "{";
    const char* names[Tuner::NK] = {"coarse", "fine", "coarse_gather", "fine_gather"};
    for (int kind = 0; kind < Tuner::NK; ++kind) {
        o += std::string(kind ? ", " : "") + "\"" + names[kind] + "\": {";
        bool first = true;
        for (int b = 0; b < Tuner::NB; ++b) {
            bool any = false;
            for (int cd = 0; cd < Tuner::NC; ++cd) any = any || c->tune.cnt[kind][b][cd];
            if (!any) continue;
            char t[256];
            const int bc = c->tune.best(kind, b);
            if (kind < 2)
                std::snprintf(t, sizeof(t), "%s\"%u-%u\": {\"best_waves\": %d", first ? "" : ", ", 1u << b,
                              (2u << b) - 1u, Tuner::tw(bc));
            else
                std::snprintf(t, sizeof(t), "%s\"%u-%u\": {\"best_kib\": %d", first ? "" : ", ", 1u << b,
                              (2u << b) - 1u, Tuner::kib(bc));
            o += t;
            for (int cd = 0; cd < Tuner::NC; ++cd)
                if (c->tune.cnt[kind][b][cd]) {
                    if (kind < 2)
                        std::snprintf(t, sizeof(t), ", \"us_%dw\": %.2f", Tuner::tw(cd),
                                      1e3 * c->tune.ms[kind][b][cd] / c->tune.cnt[kind][b][cd]);
                    else
                        std::snprintf(t, sizeof(t), ", \"us_%dkib\": %.2f", Tuner::kib(cd),
                                      1e3 * c->tune.ms[kind][b][cd] / c->tune.cnt[kind][b][cd]);
                    o += t;
                }
            o += "}";
            first = false;
        }
        o += "}";
    }
    char t[256];
    std::snprintf(t, sizeof(t),
                  ", \"chunk_graphs\": {\"on\": %d, \"captured\": %llu, \"replayed\": %llu, "
                  "\"phase_skipping_chunks\": %llu, \"pauses\": %llu}",
                  c->graphs_on ? 1 : 0, (unsigned long long)c->graph_captures,
                  (unsigned long long)c->graph_replays, (unsigned long long)c->skip_chunks,
                  (unsigned long long)c->pauses);
    o += t;
    o += "}";
    std::snprintf(buf, (size_t)cap, "%s", o.c_str());
    return SM2D_OK;
}

int sm2d_get_timing(sm2d_ctx* c, sm2d_timing* out) {
    if (!c || !out || out->struct_size != sizeof(sm2d_timing)) return err("sm2d_get_timing: struct_size");
    *out = c->timing;
    return SM2D_OK;
}

int sm2d_math_eval(int which, const float* x, float* y, int n) {
    if (n <= 0) return SM2D_OK;
    float *dx = nullptr, *dy = nullptr;
    HIPCHK(hipMalloc((void**)&dx, sizeof(float) * n));
    HIPCHK(hipMalloc((void**)&dy, sizeof(float) * n));
    HIPCHK(hipMemcpy(dx, x, sizeof(float) * n, hipMemcpyHostToDevice));
    k_math_eval<<<(n + 255) / 256, 256>>>(which, dx, dy, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost));
    (void)hipFree(dx);
    (void)hipFree(dy);
    return SM2D_OK;
}

}  // extern "C"
