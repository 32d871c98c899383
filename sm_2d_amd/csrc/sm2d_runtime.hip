// sm2d_runtime.hip — context, per-iteration orchestration and the C ABI of libsm2d_hip.so.
//
// One context = one problem grid on one device, with its own HIP stream and all device memory
// (persistent across spots).  The iteration mirrors run_k1k6_pipeline_transport
// (src/cuda/k1k6_pipeline.cu:1540-2038):
//   reset psi_out -> K1 mask -> K2 compaction -> [coarse K3 -> K4] -> [fine K3 -> K4] -> K5 -> K6
// without the reference's 60 host-synchronous calls: one host read-back per iteration (list
// counts + K5 report) in the same stream.
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

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
    hipError_t alloc(size_t count) {
        n = count;
        return hipMalloc((void**)&p, sizeof(T) * (count ? count : 1));
    }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
};

enum LayoutId { L_REF = 0, L_CFG = 1 };

}  // namespace

struct sm2d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    sm2d_grid g{};
    LayoutId layout = L_REF;
    int LB = 256, N = 0;
    uint32_t pcap = 0, lcap = 0;
    bool tables = false;
    // tables
    DBuf<float> te, ee, R, S, lE, lR, lS, feA, feB, rho, x0r;
    std::vector<float> h_ee;
    // phase space (ping-pong)
    DBuf<uint32_t> bid[2], ns[2], sref[2], pcount[2];
    DBuf<float> pool[2];
    int in = 0;
    // buckets
    DBuf<uint32_t> bbid, bn, bsref, bpcount;
    DBuf<float> bpool;
    // lists and tags
    DBuf<uint32_t> maskv, ltag, lpos, dtag, act, crs, dlist, counts /*0 nact 1 ncrs 2 ndest*/;
    DBuf<uint8_t> flag;
    DBuf<uint32_t> bcount;
    uint32_t epoch = 1;
    // tallies
    DBuf<double> edep, E_cut, E_nuc, E_bnd, win, ein, k4w, k4e, acc_c, acc_f, part, k5v;
    DBuf<float> w_cut, w_nuc, w_bnd;
    DBuf<uint32_t> flags;
    // injection
    DBuf<InjSample> samples;
    DBuf<double> injacc;
    // host source accounting of the current spot
    double src[7] = {0, 0, 0, 0, 0, 0, 0};  // W in/out/slot, E in/out/slot, E repr loss

    Tabs tabs() const {
        Tabs t;
        t.Nx = g.Nx; t.Nz = g.Nz; t.N = N; t.NT = g.N_theta; t.NE = g.N_E;
        t.dx = g.dx; t.dz = g.dz; t.x_min = g.x_min; t.z_min = g.z_min;
        t.te = te.p; t.ee = ee.p; t.R = R.p; t.lR = lR.p; t.lE = lE.p;
        t.feA = feA.p; t.feB = feB.p; t.rho = rho.p; t.x0r = x0r.p;
        return t;
    }
    Psi psi(int k) const {
        Psi p;
        p.bid = bid[k].p; p.ns = ns[k].p; p.sref = sref[k].p; p.pool = pool[k].p;
        p.pcount = pcount[k].p; p.pcap = pcap;
        return p;
    }
    Bkt bkt() const {
        Bkt b;
        b.bid = bbid.p; b.n = bn.p; b.sref = bsref.p; b.pool = bpool.p; b.pcount = bpcount.p; b.pcap = pcap;
        return b;
    }
    Tally tally() const {
        Tally t;
        t.edep = edep.p; t.w_cut = w_cut.p; t.E_cut = E_cut.p; t.w_nuc = w_nuc.p; t.E_nuc = E_nuc.p;
        t.w_bnd = w_bnd.p; t.E_bnd = E_bnd.p;
        return t;
    }
    int blocks256() const { return (N + 255) / 256; }
};

namespace {

int grid_blocks(uint32_t n, int per_block, int cap = 4096) {
    int b = (int)((n + per_block - 1) / per_block);
    return b < 1 ? 1 : (b > cap ? cap : b);
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
    p.E_on = t->E_fine_on;
    p.E_off = t->E_fine_off;
    p.wmin = t->weight_active_min;
    p.step_coarse = t->step_coarse;
    p.sx0 = t->sigma_x_initial;
    p.b_on = b_threshold(c, t->E_fine_on);
    p.b_off = b_threshold(c, t->E_fine_off);
    p.rule = t->step_rule;
    p.strag = t->enable_straggling;
    p.nuc = t->enable_nuclear;
    return p;
}

template <class LY>
void launch_clear_out(sm2d_ctx* c) {
    k_clear<<<grid_blocks(c->N, 256, 2048), 256, 0, c->stream>>>(c->psi(1 - c->in), c->N);
}

// K1 + K2: lists of the current psi_in, counts copied to host.
template <class LY>
int mask_compact(sm2d_ctx* c, const Tp& P, uint32_t ep, uint32_t* h_counts) {
    const int nb = c->blocks256();
    k_mask<LY><<<nb, 256, 0, c->stream>>>(c->tabs(), P, c->psi(c->in), c->maskv.p, ep, c->flag.p,
                                           c->bcount.p, c->win.p, c->ein.p);
    k_scan<<<1, 1024, 0, c->stream>>>(c->bcount.p, nb, c->counts.p);
    k_emit<<<nb, 256, 0, c->stream>>>(c->N, c->flag.p, c->bcount.p, c->act.p, c->crs.p, c->ltag.p, c->lpos.p,
                                      2 * ep, 2 * ep + 1);
    HIPCHK(hipGetLastError());
    if (h_counts) {
        HIPCHK(hipMemcpyAsync(h_counts, c->counts.p, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return SM2D_OK;
}

template <class LY>
void run_phase(sm2d_ctx* c, const Tp& P, bool fine, uint32_t ep, uint32_t n_host) {
    if (n_host == 0) return;
    const uint32_t tag = fine ? 2 * ep + 1 : 2 * ep;
    const uint32_t* list = fine ? c->act.p : c->crs.p;
    const uint32_t* nl = c->counts.p + (fine ? 0 : 1);
    double* acc = fine ? c->acc_f.p : c->acc_c.p;
    Psi in = c->psi(c->in), out = c->psi(1 - c->in);
    k_fill_u32<<<1, 64, 0, c->stream>>>(c->bpcount.p, 0u, 1);
    k_fill_u32<<<1, 64, 0, c->stream>>>(c->counts.p + 2, 0u, 1);
    const int tb = grid_blocks(n_host, 4, 8192);  // 4 waves (cells) per block
    if (fine)
        k_transport<LY, true><<<tb, 256, 0, c->stream>>>(c->tabs(), P, in, out, c->bkt(), list, nl, c->ltag.p,
                                                          tag, c->tally(), acc, c->flags.p);
    else
        k_transport<LY, false><<<tb, 256, 0, c->stream>>>(c->tabs(), P, in, out, c->bkt(), list, nl, c->ltag.p,
                                                           tag, c->tally(), acc, c->flags.p);
    k_dest<<<grid_blocks(n_host, 256, 1024), 256, 0, c->stream>>>(list, nl, c->g.Nx, c->g.Nz, c->dtag.p, tag,
                                                                   c->dlist.p, c->counts.p + 2, c->lcap, c->flags.p);
    k_gather<LY><<<grid_blocks(4 * n_host, 4, 8192), 256, 0, c->stream>>>(
        c->tabs(), out, c->bkt(), c->dlist.p, c->counts.p + 2, c->ltag.p, c->lpos.p, tag, c->k4w.p, c->k4e.p,
        c->flags.p);
}

int check_flags(sm2d_ctx* c) {
    uint32_t f[NFLAG];
    HIPCHK(hipMemcpyAsync(f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (f[F_PSI_FULL]) return err("phase-space payload pool full (raise pool capacity)", SM2D_ERR_CAPACITY);
    if (f[F_BKT_FULL]) return err("bucket payload pool full (raise pool capacity)", SM2D_ERR_CAPACITY);
    if (f[F_DEST_FULL]) return err("destination list full", SM2D_ERR_CAPACITY);
    return SM2D_OK;
}

// K5 of the iteration just transported (before the swap).  v: NR1 cell sums, NACC phase sums,
// 2 closure maxima.
template <class LY>
int audit(sm2d_ctx* c, std::vector<double>& v) {
    const int nb = c->blocks256();
    k_audit_cells<LY><<<nb, 256, 0, c->stream>>>(c->tabs(), c->psi(1 - c->in), c->win.p, c->ein.p, c->k4w.p,
                                                  c->k4e.p, c->part.p);
    k_audit_final<<<1, 256, 0, c->stream>>>(c->part.p, nb, c->acc_c.p, c->counts.p + 1, c->acc_f.p,
                                            c->counts.p + 0, c->k5v.p);
    HIPCHK(hipGetLastError());
    v.assign(NR1 + NACC + 2, 0.0);
    HIPCHK(hipMemcpyAsync(v.data(), c->k5v.p, sizeof(double) * v.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return SM2D_OK;
}

void fill_report(const std::vector<double>& v, double Wso, double Wss, double Eso, double Ess, int src,
                 sm2d_k5_report* r) {
    std::memset(r, 0, sizeof(*r));
    r->struct_size = sizeof(*r);
    r->W_in = v[R_WIN];
    r->E_in = v[R_EIN];
    r->W_out = v[R_WOUT];
    r->E_out = v[R_EOUT];
    const double* a = v.data() + NR1;
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

template <class LY>
int run_iters(sm2d_ctx* c, const sm2d_transport* tp, int n_iter, int src_terms, sm2d_k5_report* last,
              sm2d_audit* au) {
    Tp P = make_tp(c, tp);
    std::vector<double> v;
    for (int it = 0; it < n_iter; ++it) {
        const uint32_t ep = ++c->epoch;
        launch_clear_out<LY>(c);
        uint32_t cnt[2] = {0, 0};
        int rc = mask_compact<LY>(c, P, ep, cnt);
        if (rc) return rc;
        if (au) au->iterations++;
        if (cnt[0] == 0 && cnt[1] == 0) {
            if (au) au->terminated = 1;
            return SM2D_OK;
        }
        if (cnt[0] > c->lcap || cnt[1] > c->lcap) return err("list capacity", SM2D_ERR_CAPACITY);
        run_phase<LY>(c, P, false, ep, cnt[1]);
        run_phase<LY>(c, P, true, ep, cnt[0]);
        HIPCHK(hipGetLastError());
        rc = audit<LY>(c, v);
        if (rc) return rc;
        rc = check_flags(c);
        if (rc) return rc;
        sm2d_k5_report r;
        const bool first = src_terms && it == 0;
        fill_report(v, c->src[1], c->src[2], c->src[4], c->src[5], first ? 1 : 0, &r);
        if (last) *last = r;
        if (au) {
            const double* a = v.data() + NR1;
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
            au->processed_components += (uint64_t)a[A_COMPS];
            au->processed_cells += (uint64_t)cnt[0] + cnt[1];
        }
        if (tp->fail_fast_on_audit && (!r.W_pass || !r.E_pass))
            return err("K5 audit fail-fast", SM2D_ERR_AUDIT);
        c->in = 1 - c->in;  // K6
    }
    return SM2D_OK;
}

template <class LY>
int inject(sm2d_ctx* c, const sm2d_spot* sp) {
    const bool gauss = (sp->n_samples > 1) || (sp->sigma_x > 0) || (sp->sigma_theta > 0);
    int n = gauss ? sp->n_samples : 1;
    if (n <= 0) return err("n_samples must be positive");
    if (c->samples.n < (size_t)n) {
        if (c->samples.p) (void)hipFree(c->samples.p);
        c->samples.p = nullptr;
        HIPCHK(c->samples.alloc(n));
    }
    if (gauss)
        k_inject_samples<LY><<<(n + 255) / 256, 256, 0, c->stream>>>(c->tabs(), *sp, c->samples.p);
    else
        k_inject_pencil<LY><<<1, 1, 0, c->stream>>>(c->tabs(), *sp, c->samples.p);
    float w = gauss ? sp->W_total / sp->n_samples : sp->W_total;
    k_inject_apply<LY><<<1, 64, 0, c->stream>>>(c->psi(c->in), c->samples.p, n, w, c->injacc.p, gauss ? 0 : 1,
                                                sp->E0, c->flags.p);
    HIPCHK(hipGetLastError());
    double acc[6];
    HIPCHK(hipMemcpyAsync(acc, c->injacc.p, sizeof(acc), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int rc = check_flags(c);
    if (rc) return rc;
    // represented energy of psi_in (gpu_transport_wrapper.cu:125-180,443-457)
    std::vector<double> v;
    Psi pin = c->psi(c->in);
    const int nb = c->blocks256();
    k_audit_cells<LY><<<nb, 256, 0, c->stream>>>(c->tabs(), pin, c->win.p, c->ein.p, c->k4w.p, c->k4e.p,
                                                  c->part.p);
    k_fill_u32<<<1, 64, 0, c->stream>>>(c->counts.p, 0u, 2);
    k_audit_final<<<1, 256, 0, c->stream>>>(c->part.p, nb, c->acc_c.p, c->counts.p + 1, c->acc_f.p,
                                            c->counts.p + 0, c->k5v.p);
    v.assign(NR1 + NACC + 2, 0.0);
    HIPCHK(hipMemcpyAsync(v.data(), c->k5v.p, sizeof(double) * v.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double loss = acc[3] - v[R_EOUT];
    if (std::fabs(loss) < 1e-12) loss = 0.0;
    for (int k = 0; k < 6; ++k) c->src[k] = acc[k];
    c->src[6] = loss;
    return SM2D_OK;
}

int zero_state(sm2d_ctx* c) {
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipMemsetAsync(c->bid[k].p, 0xFF, sizeof(uint32_t) * c->bid[k].n, c->stream));
        HIPCHK(hipMemsetAsync(c->ns[k].p, 0, sizeof(uint32_t) * c->ns[k].n, c->stream));
        HIPCHK(hipMemsetAsync(c->pcount[k].p, 0, sizeof(uint32_t), c->stream));
    }
    HIPCHK(hipMemsetAsync(c->edep.p, 0, sizeof(double) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->E_cut.p, 0, sizeof(double) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->E_nuc.p, 0, sizeof(double) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->E_bnd.p, 0, sizeof(double) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->w_cut.p, 0, sizeof(float) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->w_nuc.p, 0, sizeof(float) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->w_bnd.p, 0, sizeof(float) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->k4w.p, 0, sizeof(double) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->k4e.p, 0, sizeof(double) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->flags.p, 0, sizeof(uint32_t) * NFLAG, c->stream));
    c->epoch += 4;  // stale mask / list / destination tags can no longer match
    c->in = 0;
    for (double& s : c->src) s = 0.0;
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

}  // namespace

extern "C" {

const char* sm2d_last_error(void) { return sm2d_host::error(); }
int sm2d_abi_version(void) { return SM2D_ABI_VERSION; }

int sm2d_ctx_create(int device, const sm2d_grid* g, sm2d_ctx** out) {
    if (!g || !out || g->struct_size != sizeof(sm2d_grid)) return err("sm2d_ctx_create: grid struct_size");
    if (g->Nx <= 0 || g->Nz <= 0 || g->N_theta <= 0 || g->N_E <= 0 || g->N_theta > 4095 * g->n_theta_local)
        return err("sm2d_ctx_create: bad grid");
    LayoutId lid;
    if (g->n_theta_local == 4 && g->n_e_local == 2 && g->n_x_sub == 8 && g->n_z_sub == 4) lid = L_REF;
    else if (g->n_theta_local == 8 && g->n_e_local == 4 && g->n_x_sub == 4 && g->n_z_sub == 1) lid = L_CFG;
    else return err("sm2d_ctx_create: local-bin layout must be (4,2,8,4) or (8,4,4,1)");
    HIPCHK(hipSetDevice(device));
    std::unique_ptr<sm2d_ctx> c(new sm2d_ctx());
    c->device = device;
    c->g = *g;
    c->layout = lid;
    c->LB = g->n_theta_local * g->n_e_local * g->n_x_sub * g->n_z_sub;
    c->N = g->Nx * g->Nz;
    const size_t N = c->N;
    // Payload pool: enough for every cell of a 2-row beam front band with all 32 slots, at
    // least 64 Ki slots, at most one slot table's worth.
    size_t pc = std::max<size_t>(65536, (size_t)g->Nx * 2 * KB);
    pc = std::min(pc, N * KB);
    c->pcap = (uint32_t)pc;
    c->lcap = (uint32_t)N;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        HIPCHK(c->bid[k].alloc(N * KB));
        HIPCHK(c->ns[k].alloc(N));
        HIPCHK(c->sref[k].alloc(N * KB));
        HIPCHK(c->pcount[k].alloc(1));
        HIPCHK(c->pool[k].alloc(pc * c->LB));
    }
    HIPCHK(c->bbid.alloc(N * 4 * KB));
    HIPCHK(c->bn.alloc(N * 4));
    HIPCHK(c->bsref.alloc(N * 4 * KB));
    HIPCHK(c->bpcount.alloc(1));
    HIPCHK(c->bpool.alloc(pc * c->LB));
    HIPCHK(c->maskv.alloc(N));
    HIPCHK(c->ltag.alloc(N));
    HIPCHK(c->lpos.alloc(N));
    HIPCHK(c->dtag.alloc(N));
    HIPCHK(c->act.alloc(N));
    HIPCHK(c->crs.alloc(N));
    HIPCHK(c->dlist.alloc(N));
    HIPCHK(c->counts.alloc(4));
    HIPCHK(c->flag.alloc(N));
    HIPCHK(c->bcount.alloc(2 * c->blocks256()));
    for (DBuf<double>* b : {&c->edep, &c->E_cut, &c->E_nuc, &c->E_bnd, &c->win, &c->ein, &c->k4w, &c->k4e})
        HIPCHK(b->alloc(N));
    HIPCHK(c->acc_c.alloc(N * NACC));
    HIPCHK(c->acc_f.alloc(N * NACC));
    HIPCHK(c->part.alloc((size_t)c->blocks256() * NR1));
    HIPCHK(c->k5v.alloc(NR1 + NACC + 2));
    for (DBuf<float>* b : {&c->w_cut, &c->w_nuc, &c->w_bnd}) HIPCHK(b->alloc(N));
    HIPCHK(c->flags.alloc(NFLAG));
    HIPCHK(c->injacc.alloc(6));
    HIPCHK(c->te.alloc(g->N_theta + 1));
    HIPCHK(c->ee.alloc(g->N_E + 1));
    for (DBuf<float>* b : {&c->R, &c->S, &c->lE, &c->lR, &c->lS}) HIPCHK(b->alloc(g->N_E));
    for (DBuf<float>* b : {&c->feA, &c->feB, &c->rho, &c->x0r}) HIPCHK(b->alloc(g->Nz));
    HIPCHK(hipMemsetAsync(c->maskv.p, 0, sizeof(uint32_t) * N, c->stream));
    HIPCHK(hipMemsetAsync(c->ltag.p, 0, sizeof(uint32_t) * N, c->stream));
    HIPCHK(hipMemsetAsync(c->dtag.p, 0, sizeof(uint32_t) * N, c->stream));
    HIPCHK(hipMemsetAsync(c->feA.p, 0, sizeof(float) * g->Nz, c->stream));
    HIPCHK(hipMemsetAsync(c->feB.p, 0, sizeof(float) * g->Nz, c->stream));
    std::vector<float> ones(g->Nz, 1.0f);
    HIPCHK(hipMemcpyAsync(c->rho.p, ones.data(), sizeof(float) * g->Nz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->x0r.p, ones.data(), sizeof(float) * g->Nz, hipMemcpyHostToDevice, c->stream));
    int rc = zero_state(c.get());
    if (rc) return rc;
    *out = c.release();
    return SM2D_OK;
}

int sm2d_ctx_destroy(sm2d_ctx* c) {
    if (!c) return SM2D_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
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
    HIPCHK(hipMemcpyAsync(c->S.p, S, sizeof(float) * NE, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->lE.p, lE, sizeof(float) * NE, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->lR.p, lR, sizeof(float) * NE, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->lS.p, lS, sizeof(float) * NE, hipMemcpyHostToDevice, c->stream));
    c->h_ee.assign(ee, ee + NE + 1);
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
    rc = dispatch(c, [&](auto ly) { return inject<decltype(ly)>(c, sp); });
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
    for (size_t cl = 0; cl < N; ++cl)
        for (uint32_t s = 0; s < hn[cl]; ++s) {
            size_t k = cl * KB + s;
            if (where[k] < 0) return err("psi_load: slots of a cell must be a prefix");
            hs[k] = (uint32_t)where[k];
        }
    uint32_t used = (uint32_t)(pool.size() / LB);
    if (used > c->pcap) return err("psi_load: pool capacity", SM2D_ERR_CAPACITY);
    rc = zero_state(c);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->bid[0].p, hb.data(), sizeof(uint32_t) * hb.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ns[0].p, hn.data(), sizeof(uint32_t) * hn.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->sref[0].p, hs.data(), sizeof(uint32_t) * hs.size(), hipMemcpyHostToDevice, c->stream));
    if (!pool.empty())
        HIPCHK(hipMemcpyAsync(c->pool[0].p, pool.data(), sizeof(float) * pool.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->pcount[0].p, &used, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
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
    std::vector<uint32_t> hb(N * KB), hn(N), hs(N * KB), cnt(1);
    HIPCHK(hipMemcpyAsync(hb.data(), c->bid[k].p, sizeof(uint32_t) * hb.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hn.data(), c->ns[k].p, sizeof(uint32_t) * hn.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hs.data(), c->sref[k].p, sizeof(uint32_t) * hs.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(cnt.data(), c->pcount[k].p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    uint32_t used = std::min(cnt[0], c->pcap);
    std::vector<float> pool((size_t)used * LB);
    if (used)
        HIPCHK(hipMemcpyAsync(pool.data(), c->pool[k].p, sizeof(float) * pool.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (bids) std::memcpy(bids, hb.data(), sizeof(uint32_t) * hb.size());
    int m = 0;
    for (size_t cl = 0; cl < N; ++cl)
        for (uint32_t s = 0; s < hn[cl]; ++s) {
            const float* v = pool.data() + (size_t)hs[cl * KB + s] * LB;
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
    const uint32_t ep = ++c->epoch;
    std::vector<uint32_t> mv(c->N, 0);
    if (mask_prev)
        for (int i = 0; i < c->N; ++i) mv[i] = mask_prev[i] ? (((ep - 1u) << 1) | 1u) : 0u;
    HIPCHK(hipMemcpyAsync(c->maskv.p, mv.data(), sizeof(uint32_t) * c->N, hipMemcpyHostToDevice, c->stream));
    Tp P = make_tp(c, tp);
    uint32_t cnt[2];
    rc = dispatch(c, [&](auto ly) { return mask_compact<decltype(ly)>(c, P, ep, cnt); });
    if (rc) return rc;
    std::vector<uint32_t> a(cnt[0] ? cnt[0] : 1), k(cnt[1] ? cnt[1] : 1);
    HIPCHK(hipMemcpyAsync(mv.data(), c->maskv.p, sizeof(uint32_t) * c->N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(a.data(), c->act.p, sizeof(uint32_t) * cnt[0], hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(k.data(), c->crs.p, sizeof(uint32_t) * cnt[1], hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (mask_out)
        for (int i = 0; i < c->N; ++i) mask_out[i] = (mv[i] == ((ep << 1) | 1u)) ? 1 : 0;
    for (uint32_t i = 0; i < cnt[0]; ++i) active[i] = (int32_t)a[i];
    for (uint32_t i = 0; i < cnt[1]; ++i) coarse[i] = (int32_t)k[i];
    *n_active = (int)cnt[0];
    *n_coarse = (int)cnt[1];
    return SM2D_OK;
}

static int download_tallies(sm2d_ctx* c, double* t) {
    const size_t N = c->N;
    std::vector<float> f(N);
    HIPCHK(hipMemcpyAsync(t + 0 * N, c->edep.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(t + 2 * N, c->E_cut.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(t + 4 * N, c->E_nuc.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(t + 6 * N, c->E_bnd.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    const float* src[3] = {c->w_cut.p, c->w_nuc.p, c->w_bnd.p};
    const int dst[3] = {1, 3, 5};
    for (int k = 0; k < 3; ++k) {
        HIPCHK(hipMemcpyAsync(f.data(), src[k], sizeof(float) * N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < N; ++i) t[dst[k] * N + i] = f[i];
    }
    HIPCHK(hipStreamSynchronize(c->stream));
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
    rc = dispatch(c, [&](auto ly) { return run_iters<decltype(ly)>(c, tp, n_iter, src_terms, last, au); });
    if (rc) return rc;
    if (au) {
        uint32_t f[NFLAG];
        HIPCHK(hipMemcpy(f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost));
        au->unsupported_events = f[F_UNSUPPORTED];
    }
    if (tallies) return download_tallies(c, tallies);
    return SM2D_OK;
}

int sm2d_k5_audit(sm2d_ctx* c, const double* td, double Wso, double Wss, double Eso, double Ess, double Wd,
                  double Ed, int src, sm2d_k5_report* out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!td || !out) return err("sm2d_k5_audit: null argument");
    HIPCHK(hipSetDevice(c->device));
    // psi_in / psi_out totals on the device, tally deltas summed on the host in cell order
    auto totals = [&](int k, double& W, double& E) -> int {
        const int nb = c->blocks256();
        Psi p = c->psi(k);
        int r2 = dispatch(c, [&](auto ly) {
            k_audit_cells<decltype(ly)><<<nb, 256, 0, c->stream>>>(c->tabs(), p, c->win.p, c->ein.p, c->k4w.p,
                                                                   c->k4e.p, c->part.p);
            return 0;
        });
        (void)r2;
        k_fill_u32<<<1, 64, 0, c->stream>>>(c->counts.p, 0u, 2);
        k_audit_final<<<1, 256, 0, c->stream>>>(c->part.p, nb, c->acc_c.p, c->counts.p + 1, c->acc_f.p,
                                                c->counts.p, c->k5v.p);
        std::vector<double> v(NR1 + NACC + 2);
        HIPCHK(hipMemcpyAsync(v.data(), c->k5v.p, sizeof(double) * v.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        W = v[R_WOUT];
        E = v[R_EOUT];
        return SM2D_OK;
    };
    HIPCHK(hipMemsetAsync(c->win.p, 0, sizeof(double) * c->N, c->stream));
    HIPCHK(hipMemsetAsync(c->ein.p, 0, sizeof(double) * c->N, c->stream));
    double Wi, Ei, Wo, Eo;
    rc = totals(c->in, Wi, Ei);
    if (rc) return rc;
    rc = totals(1 - c->in, Wo, Eo);
    if (rc) return rc;
    const size_t N = c->N;
    std::vector<double> v(NR1 + NACC + 2, 0.0);
    double* a = v.data() + NR1;
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
    rc = sm2d_inject(c, sp, au);
    if (rc) return rc;
    sm2d_transport t2 = *tp;
    t2.sigma_x_initial = sp->sigma_x;
    sm2d_audit it;
    std::memset(&it, 0, sizeof(it));
    rc = dispatch(c, [&](auto ly) { return run_iters<decltype(ly)>(c, &t2, tp->max_iterations, 1, nullptr, &it); });
    if (rc) return rc;
    // keep the source block, merge the transport block
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
    au->processed_cells = it.processed_cells;
    au->processed_components = it.processed_components;
    std::vector<double> t(7 * (size_t)c->N);
    rc = download_tallies(c, t.data());
    if (rc) return rc;
    const size_t N = c->N;
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 7; ++k)
        for (size_t i = 0; i < N; ++i) s[k] += t[k * N + i];
    au->E_dep = s[0];
    au->W_cutoff = s[1];
    au->E_cutoff = s[2];
    au->W_nuclear = s[3];
    au->E_nuclear = s[4];
    au->W_boundary = s[5];
    au->E_boundary = s[6];
    if (!au->terminated) {
        // phase space left when max_iterations stopped the loop (not accounted by the reference)
        std::vector<double> v;
        const int nb = c->blocks256();
        int r2 = dispatch(c, [&](auto ly) {
            k_audit_cells<decltype(ly)><<<nb, 256, 0, c->stream>>>(c->tabs(), c->psi(c->in), c->win.p, c->ein.p,
                                                                   c->k4w.p, c->k4e.p, c->part.p);
            return 0;
        });
        (void)r2;
        k_fill_u32<<<1, 64, 0, c->stream>>>(c->counts.p, 0u, 2);
        k_audit_final<<<1, 256, 0, c->stream>>>(c->part.p, nb, c->acc_c.p, c->counts.p + 1, c->acc_f.p,
                                                c->counts.p, c->k5v.p);
        v.assign(NR1 + NACC + 2, 0.0);
        HIPCHK(hipMemcpyAsync(v.data(), c->k5v.p, sizeof(double) * v.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        au->W_leftover = v[R_WOUT];
        au->E_leftover = v[R_EOUT];
    }
    uint32_t f[NFLAG];
    HIPCHK(hipMemcpy(f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost));
    au->unsupported_events = f[F_UNSUPPORTED];
    if (edep_out) std::memcpy(edep_out, t.data(), sizeof(double) * N);
    if (f[F_UNSUPPORTED])
        return err("lateral emission into another listed cell's bucket is not supported on the GPU path",
                   SM2D_ERR_UNSUPPORTED);
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
