// sm2d_surrogate.hip — host side of the learned part: the MLP object (FP32 master parameters,
// Adam state, BF16 copies), its forward / backward passes built from the matrix-core GEMM of
// sm2d_mlp.hip, training on samples of the plain physics generated on the GPU, and the
// surrogate table that the fine transport reads.  Included by sm2d_runtime.hip (one
// translation unit), after the context definition.

namespace {

using namespace sm2d::mlp;

struct MlpWork {
    int Bmax = 0;
    DBuf<uint16_t> X, A1, A2, A3, dZb, dZb4;
    DBuf<float> Z1, Z2, Z3, Z4, dZ4, dA, dZ, P, lossv, lossum, labels;
};

}  // namespace

struct sm2d_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    DBuf<float> p, g, m1, m2;
    DBuf<uint16_t> wb;
    Dict dict{0.0f, 1.0f, 0.0f, 1.0f};
    int has_dict = 0;
    int step = 0;
    MlpWork w;
};

namespace {

constexpr int KSPLIT_W = 16;  // K splits of the weight-gradient GEMMs (reduction over the batch)

int mlp_ensure(sm2d_mlp* M, int Bn) {
    if (Bn <= M->w.Bmax) return SM2D_OK;
    MlpWork& w = M->w;
    w.Bmax = Bn;
    HIPCHK(w.X.alloc((size_t)Bn * KP0));
    for (DBuf<uint16_t>* b : {&w.A1, &w.A2, &w.A3, &w.dZb}) HIPCHK(b->alloc((size_t)Bn * NH));
    HIPCHK(w.dZb4.alloc((size_t)Bn * NOP));
    for (DBuf<float>* b : {&w.Z1, &w.Z2, &w.Z3, &w.dA, &w.dZ}) HIPCHK(b->alloc((size_t)Bn * NH));
    HIPCHK(w.Z4.alloc((size_t)Bn * NOP));
    HIPCHK(w.dZ4.alloc((size_t)Bn * NOP));
    HIPCHK(w.P.alloc(std::max<size_t>((size_t)Bn * NH, (size_t)KSPLIT_W * NH * NH)));
    HIPCHK(w.lossv.alloc(Bn));
    HIPCHK(w.lossum.alloc(1));
    HIPCHK(w.labels.alloc((size_t)Bn * 3));
    return SM2D_OK;
}

// P = A . B on the matrix cores, then the epilogue into Z / act.
int gemm(hipStream_t st, const uint16_t* A, int sam, int sak, const uint16_t* B, int sbk, int sbn, int Mr, int N, int K,
         int ksplit, float* P, int mvalid, int nvalid, const float* bias, int mode, float* Z, int ldz, uint16_t* act,
         int lda) {
    dim3 grid((Mr + 63) / 64, (N + 63) / 64, ksplit);
    k_gemm_bf16<<<grid, 256, 0, st>>>(A, sam, sak, B, sbk, sbn, Mr, N, K, ksplit, P);
    const size_t tot = (size_t)Mr * N;
    k_epilogue<<<(int)std::min<size_t>((tot + 255) / 256, 4096), 256, 0, st>>>(P, ksplit, Mr, N, mvalid, nvalid, bias,
                                                                             mode, Z, ldz, act, lda);
    HIPCHK(hipGetLastError());
    return SM2D_OK;
}

// Forward of Bn samples already in w.X (Bn multiple of 64): Z1..Z4, A1..A3.
int mlp_forward(sm2d_mlp* M, int Bn) {
    MlpWork& w = M->w;
    const float* p = M->p.p;
    const uint16_t* wb = M->wb.p;
    hipStream_t st = M->stream;
    int rc;
    if ((rc = gemm(st, w.X.p, KP0, 1, wb + BOFF_W0, 1, KP0, Bn, NH, KP0, 1, w.P.p, Bn, NH, p + OFF_B0, 1, w.Z1.p, NH,
                   w.A1.p, NH)))
        return rc;
    if ((rc = gemm(st, w.A1.p, NH, 1, wb + BOFF_W1, 1, NH, Bn, NH, NH, 1, w.P.p, Bn, NH, p + OFF_B1, 1, w.Z2.p, NH,
                   w.A2.p, NH)))
        return rc;
    if ((rc = gemm(st, w.A2.p, NH, 1, wb + BOFF_W2, 1, NH, Bn, NH, NH, 1, w.P.p, Bn, NH, p + OFF_B2, 1, w.Z3.p, NH,
                   w.A3.p, NH)))
        return rc;
    return gemm(st, w.A3.p, NH, 1, wb + BOFF_W3, 1, NH, Bn, NOP, NH, 1, w.P.p, Bn, NOUT, p + OFF_B3, 0, w.Z4.p, NOP,
                nullptr, 0);
}

// Loss and gradients (into M->g) of Bn samples in w.X with labels in w.labels.
int mlp_backward(sm2d_mlp* M, int Bn, float w_att, float w_mom) {
    MlpWork& w = M->w;
    float* g = M->g.p;
    const uint16_t* wb = M->wb.p;
    hipStream_t st = M->stream;
    const int gb = std::min((Bn + 255) / 256, 4096);
    int rc;
    k_loss<<<gb, 256, 0, st>>>(w.Z4.p, w.labels.p, Bn, M->dict, w_att, w_mom, w.dZ4.p, w.dZb4.p, w.lossv.p);
    k_sum<<<1, 256, 0, st>>>(w.lossv.p, Bn, w.lossum.p);
    // layer 3: dW3 = dZ4^T A3, db3, dA3 = dZ4 W3
    if ((rc = gemm(st, w.dZb4.p, 1, NOP, w.A3.p, NH, 1, NOP, NH, Bn, KSPLIT_W, w.P.p, NOUT, NH, nullptr, 0,
                   g + OFF_W3, NH, nullptr, 0)))
        return rc;
    k_colsum<<<1, 64 * COLSUM_WAVES, 0, st>>>(w.dZ4.p, Bn, NOP, NOUT, g + OFF_B3);
    if ((rc = gemm(st, w.dZb4.p, NOP, 1, wb + BOFF_W3, NH, 1, Bn, NH, NOP, 1, w.P.p, Bn, NH, nullptr, 0, w.dA.p, NH,
                   nullptr, 0)))
        return rc;
    const size_t nh = (size_t)Bn * NH;
    const int ge = (int)std::min<size_t>((nh + 255) / 256, 8192);
    struct L {
        const float* Z;
        const uint16_t* Ain;
        int ldin, kin;
        int offW, offB, boffW;
    };
    // layers 2, 1, 0 (input activations A2, A1, X)
    const L layers[3] = {{w.Z3.p, w.A2.p, NH, NH, OFF_W2, OFF_B2, BOFF_W2},
                         {w.Z2.p, w.A1.p, NH, NH, OFF_W1, OFF_B1, BOFF_W1},
                         {w.Z1.p, w.X.p, KP0, KP0, OFF_W0, OFF_B0, BOFF_W0}};
    for (int li = 0; li < 3; ++li) {
        const L& y = layers[li];
        k_gelu_back<<<ge, 256, 0, st>>>(w.dA.p, y.Z, nh, w.dZ.p, w.dZb.p);
        // dW = dZ^T Ain ([NH][kin]; the first layer keeps only NIN input columns)
        const int nval = (li == 2) ? NIN : y.kin;
        if ((rc = gemm(st, w.dZb.p, 1, NH, y.Ain, y.ldin, 1, NH, y.kin, Bn, KSPLIT_W, w.P.p, NH, nval, nullptr, 0,
                       g + y.offW, nval, nullptr, 0)))
            return rc;
        k_colsum<<<(NH + 63) / 64, 64 * COLSUM_WAVES, 0, st>>>(w.dZ.p, Bn, NH, NH, g + y.offB);
        if (li < 2) {  // dA of the previous layer = dZ W
            if ((rc = gemm(st, w.dZb.p, NH, 1, wb + (li == 0 ? BOFF_W2 : BOFF_W1), NH, 1, Bn, NH, NH, 1, w.P.p, Bn, NH,
                           nullptr, 0, w.dA.p, NH, nullptr, 0)))
                return rc;
        }
    }
    HIPCHK(hipGetLastError());
    return SM2D_OK;
}

int mlp_pack(sm2d_mlp* M) {
    k_pack_bf16<<<(NBW + 255) / 256, 256, 0, M->stream>>>(M->p.p, M->wb.p);
    HIPCHK(hipGetLastError());
    return SM2D_OK;
}

int mlp_upload_x(sm2d_mlp* M, const float* x, int B, int Bp) {
    std::vector<uint16_t> hx((size_t)Bp * KP0, 0);
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < NIN; ++k) hx[(size_t)b * KP0 + k] = f2bf(x[(size_t)b * NIN + k]);
    HIPCHK(hipMemcpyAsync(M->w.X.p, hx.data(), sizeof(uint16_t) * hx.size(), hipMemcpyHostToDevice, M->stream));
    return SM2D_OK;
}

// Compares the attached table with the plain physics at its nodes and sends every node whose
// learned ln sigma_E or ln sigma_theta differs by more than `tol`, or whose attenuation exponent
// a differs by more than tol x (a + 1e-4 / mm x dR), back to the plain physics (NaN entries).
// stats4 = (nodes checked, nodes flagged, largest deviation, RMS deviation).
int surrogate_validate(sm2d_ctx* c, float tol, float* stats4) {
    const int Nz = c->g.Nz, NE = c->g.N_E, nmat = c->surr_nmat, NJ = c->surr_nodes;
    std::vector<int> rows(Nz);
    std::vector<float> rho(Nz), x0r(Nz);
    HIPCHK(hipMemcpyAsync(rows.data(), c->rowmat.p, sizeof(int) * Nz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(rho.data(), c->rho.p, sizeof(float) * Nz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(x0r.data(), c->x0r.p, sizeof(float) * Nz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<float> mr(nmat, 1.0f), mx(nmat, 1.0f);
    std::vector<uint8_t> used(nmat, 0);
    for (int z = 0; z < Nz; ++z)
        if (rows[z] >= 0 && !used[rows[z]]) {
            used[rows[z]] = 1;
            mr[rows[z]] = rho[z];
            mx[rows[z]] = x0r[z];
        }
    const size_t total = (size_t)nmat * NE * NJ;
    DBuf<float> dmr, dmx, plain;
    DBuf<uint8_t> valid;
    HIPCHK(dmr.alloc(nmat));
    HIPCHK(dmx.alloc(nmat));
    HIPCHK(plain.alloc(total * 3));
    HIPCHK(valid.alloc(total));
    HIPCHK(hipMemcpyAsync(dmr.p, mr.data(), sizeof(float) * nmat, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dmx.p, mx.data(), sizeof(float) * nmat, hipMemcpyHostToDevice, c->stream));
    k_table_plain<<<std::min<size_t>((total + 255) / 256, 8192), 256, 0, c->stream>>>(
        c->tabs(), dmr.p, dmx.p, nmat, NJ, plain.p, valid.p);
    HIPCHK(hipGetLastError());
    std::vector<float> tab(total * 3), pl(total * 3);
    std::vector<uint8_t> hv(total);
    HIPCHK(hipMemcpyAsync(tab.data(), c->surr.p, sizeof(float) * tab.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(pl.data(), plain.p, sizeof(float) * pl.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hv.data(), valid.p, total, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double n = 0, flagged = 0, mx_dev = 0, ss = 0;
    for (size_t i = 0; i < total; ++i) {
        const int m = (int)(i / ((size_t)NE * NJ));
        if (!used[m] || !hv[i]) continue;
        float* t = &tab[3 * i];
        if (!(t[0] == t[0] && t[1] == t[1] && t[2] == t[2])) continue;  // already plain
        const double d = std::max(std::fabs((double)t[0] - pl[3 * i]), std::fabs((double)t[1] - pl[3 * i + 1]));
        // attenuation: relative error, floored at a cross-section of 1e-4 / mm (sigma_n vanishes
        // below 5 MeV, where any small exponent is harmless)
        const int j = (int)(i % (size_t)NJ);
        const double dR = std::exp(std::log(1e-4) + (std::log(10.0) - std::log(1e-4)) * j / (double)(NJ - 1));
        const double da = std::fabs((double)std::max(t[2], 0.0f) - pl[3 * i + 2]) / (pl[3 * i + 2] + 1e-4 * dR);
        const double dev = std::max(d, da);
        n += 1;
        ss += dev * dev;
        mx_dev = std::max(mx_dev, dev);
        if (tol > 0.0f && dev > tol) {
            t[0] = t[1] = t[2] = NAN;
            flagged += 1;
        }
    }
    if (flagged > 0) {
        HIPCHK(hipMemcpyAsync(c->surr.p, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->surr_stats[0] = (float)n;
    c->surr_stats[1] = (float)flagged;
    c->surr_stats[2] = (float)mx_dev;
    c->surr_stats[3] = n > 0 ? (float)std::sqrt(ss / n) : 0.0f;
    if (stats4)
        for (int k = 0; k < 4; ++k) stats4[k] = c->surr_stats[k];
    return SM2D_OK;
}

}  // namespace

extern "C" {

int sm2d_mlp_create(int device, uint32_t seed, sm2d_mlp** out) {
    if (!out) return err("sm2d_mlp_create: null argument");
    HIPCHK(hipSetDevice(device));
    std::unique_ptr<sm2d_mlp> M(new sm2d_mlp());
    M->device = device;
    HIPCHK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
    for (DBuf<float>* b : {&M->p, &M->g, &M->m1, &M->m2}) HIPCHK(b->alloc(NPARAM));
    HIPCHK(M->wb.alloc(NBW));
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and zero biases from a seeded LCG
    std::vector<float> h(NPARAM, 0.0f);
    uint64_t s = 0x2545F4914F6CDD1Dull ^ seed;
    auto uni = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (float)((s >> 40) & 0xFFFFFF) / 16777216.0f * 2.0f - 1.0f;
    };
    const int offs[4] = {OFF_W0, OFF_W1, OFF_W2, OFF_W3}, fan[4] = {NIN, NH, NH, NH}, rows[4] = {NH, NH, NH, NOUT};
    for (int L = 0; L < 4; ++L) {
        const float a = 1.0f / std::sqrt((float)fan[L]);
        for (int i = 0; i < rows[L] * fan[L]; ++i) h[offs[L] + i] = a * uni();
    }
    HIPCHK(hipMemcpyAsync(M->p.p, h.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, M->stream));
    HIPCHK(hipMemsetAsync(M->m1.p, 0, sizeof(float) * NPARAM, M->stream));
    HIPCHK(hipMemsetAsync(M->m2.p, 0, sizeof(float) * NPARAM, M->stream));
    int rc = mlp_pack(M.get());
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(M->stream));
    *out = M.release();
    return SM2D_OK;
}

int sm2d_mlp_destroy(sm2d_mlp* M) {
    if (!M) return SM2D_OK;
    (void)hipSetDevice(M->device);
    (void)hipStreamSynchronize(M->stream);
    hipStream_t s = M->stream;
    delete M;
    if (s) (void)hipStreamDestroy(s);
    return SM2D_OK;
}

int sm2d_mlp_param_count(void) { return NPARAM; }

int sm2d_mlp_get_params(sm2d_mlp* M, float* p, float* dict4) {
    if (!M || !p) return err("sm2d_mlp_get_params: null argument");
    HIPCHK(hipSetDevice(M->device));
    HIPCHK(hipMemcpyAsync(p, M->p.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, M->stream));
    HIPCHK(hipStreamSynchronize(M->stream));
    if (dict4) {
        dict4[0] = M->dict.lsE_lo;
        dict4[1] = M->dict.lsE_hi;
        dict4[2] = M->dict.lsT_lo;
        dict4[3] = M->dict.lsT_hi;
    }
    return SM2D_OK;
}

int sm2d_mlp_set_params(sm2d_mlp* M, const float* p, const float* dict4) {
    if (!M) return err("sm2d_mlp_set_params: null argument");
    HIPCHK(hipSetDevice(M->device));
    if (p) {
        HIPCHK(hipMemcpyAsync(M->p.p, p, sizeof(float) * NPARAM, hipMemcpyHostToDevice, M->stream));
        int rc = mlp_pack(M);
        if (rc) return rc;
    }
    if (dict4) {
        if (!(dict4[1] > dict4[0]) || !(dict4[3] > dict4[2])) return err("dictionary ranges must be increasing");
        M->dict = Dict{dict4[0], dict4[1], dict4[2], dict4[3]};
        M->has_dict = 1;
    }
    HIPCHK(hipStreamSynchronize(M->stream));
    return SM2D_OK;
}

// Raw network outputs z[B][33] of normalised inputs x[B][4].
int sm2d_mlp_forward(sm2d_mlp* M, const float* x, int B, float* z) {
    if (!M || !x || !z || B <= 0) return err("sm2d_mlp_forward: bad argument");
    HIPCHK(hipSetDevice(M->device));
    const int Bp = (B + 63) / 64 * 64;
    int rc = mlp_ensure(M, Bp);
    if (rc) return rc;
    if ((rc = mlp_upload_x(M, x, B, Bp))) return rc;
    if ((rc = mlp_forward(M, Bp))) return rc;
    std::vector<float> hz((size_t)Bp * NOP);
    HIPCHK(hipMemcpyAsync(hz.data(), M->w.Z4.p, sizeof(float) * hz.size(), hipMemcpyDeviceToHost, M->stream));
    HIPCHK(hipStreamSynchronize(M->stream));
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < NOUT; ++k) z[(size_t)b * NOUT + k] = hz[(size_t)b * NOP + k];
    return SM2D_OK;
}

// Decoded outputs (ln sigma_E, ln sigma_theta, a) of normalised inputs x[B][4], evaluated by the
// in-kernel inference of the fine transport (mlp_wave_forward).
int sm2d_mlp_infer(sm2d_mlp* M, const float* x, int B, float* out3) {
    if (!M || !x || !out3 || B <= 0) return err("sm2d_mlp_infer: bad argument");
    if (!M->has_dict) return err("sm2d_mlp_infer: the network has no width dictionaries");
    HIPCHK(hipSetDevice(M->device));
    DBuf<float> dx, dout;
    HIPCHK(dx.alloc((size_t)B * 4));
    HIPCHK(dout.alloc((size_t)B * 3));
    HIPCHK(hipMemcpyAsync(dx.p, x, sizeof(float) * 4 * (size_t)B, hipMemcpyHostToDevice, M->stream));
    k_mlp_wave_probe<<<(B + 63) / 64, 64, 0, M->stream>>>(M->wb.p, M->p.p, M->dict, dx.p, B, dout.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out3, dout.p, sizeof(float) * 3 * (size_t)B, hipMemcpyDeviceToHost, M->stream));
    HIPCHK(hipStreamSynchronize(M->stream));
    return SM2D_OK;
}

// Loss and its gradient (FP32, layout of sm2d_mlp_get_params) for B samples (B multiple of 64):
// x[B][4] normalised inputs, labels[B][3] = (ln sigma_E, ln sigma_theta, ln(a + 1e-9)).
int sm2d_mlp_grad(sm2d_mlp* M, const float* x, const float* labels, int B, float w_att, float w_mom, float* loss,
                  float* grad) {
    if (!M || !x || !labels || B <= 0 || B % 64) return err("sm2d_mlp_grad: B must be a positive multiple of 64");
    if (!M->has_dict) return err("sm2d_mlp_grad: width dictionaries not set");
    HIPCHK(hipSetDevice(M->device));
    int rc = mlp_ensure(M, B);
    if (rc) return rc;
    if ((rc = mlp_upload_x(M, x, B, B))) return rc;
    HIPCHK(hipMemcpyAsync(M->w.labels.p, labels, sizeof(float) * 3 * B, hipMemcpyHostToDevice, M->stream));
    if ((rc = mlp_forward(M, B))) return rc;
    if ((rc = mlp_backward(M, B, w_att, w_mom))) return rc;
    if (grad) HIPCHK(hipMemcpyAsync(grad, M->g.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, M->stream));
    if (loss) HIPCHK(hipMemcpyAsync(loss, M->w.lossum.p, sizeof(float), hipMemcpyDeviceToHost, M->stream));
    HIPCHK(hipStreamSynchronize(M->stream));
    if (loss) *loss /= (float)B;
    return SM2D_OK;
}

// Trains on n_samples samples of the plain physics drawn with the tables of `ctx` (its R-LUT
// and energy edges).  Sets the width dictionaries from the sample range on first use.
// loss_hist[i] receives the mean batch loss of the i-th block of steps (n_hist blocks).
int sm2d_mlp_train(sm2d_mlp* M, sm2d_ctx* c, const sm2d_mlp_train_cfg* cfg, float* loss_hist, int n_hist) {
    if (!M || !c || !cfg || cfg->struct_size != sizeof(sm2d_mlp_train_cfg)) return err("sm2d_mlp_train: bad argument");
    if (!c->tables) return err("sm2d_mlp_train: context tables not uploaded");
    const int B = cfg->batch;
    if (B <= 0 || B % 64 || cfg->n_samples < B || cfg->steps <= 0) return err("sm2d_mlp_train: bad sizes");
    HIPCHK(hipSetDevice(M->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    int rc = mlp_ensure(M, B);
    if (rc) return rc;
    const int n = cfg->n_samples / B * B;
    DBuf<uint16_t> TX;
    DBuf<float> TL;
    HIPCHK(TX.alloc((size_t)n * KP0));
    HIPCHK(TL.alloc((size_t)n * 3));
    k_gen_samples<<<std::min((n + 255) / 256, 8192), 256, 0, M->stream>>>(c->tabs(), cfg->seed, n, TX.p, TL.p,
                                                                          nullptr);
    HIPCHK(hipGetLastError());
    if (!M->has_dict) {
        std::vector<float> hl((size_t)n * 3);
        HIPCHK(hipMemcpyAsync(hl.data(), TL.p, sizeof(float) * hl.size(), hipMemcpyDeviceToHost, M->stream));
        HIPCHK(hipStreamSynchronize(M->stream));
        float e0 = 1e30f, e1 = -1e30f, t0 = 1e30f, t1 = -1e30f;
        for (int i = 0; i < n; ++i) {
            e0 = std::min(e0, hl[3 * i]);
            e1 = std::max(e1, hl[3 * i]);
            t0 = std::min(t0, hl[3 * i + 1]);
            t1 = std::max(t1, hl[3 * i + 1]);
        }
        M->dict = Dict{e0, e1, t0, t1};
        M->has_dict = 1;
    }
    const int nb = n / B;
    const int per = std::max(1, cfg->steps / std::max(1, n_hist));
    std::vector<float> lsum(std::max(1, n_hist), 0.0f);
    std::vector<int> lcnt(std::max(1, n_hist), 0);
    for (int s = 0; s < cfg->steps; ++s) {
        const int bidx = s % nb;
        HIPCHK(hipMemcpyAsync(M->w.X.p, TX.p + (size_t)bidx * B * KP0, sizeof(uint16_t) * B * KP0,
                              hipMemcpyDeviceToDevice, M->stream));
        HIPCHK(hipMemcpyAsync(M->w.labels.p, TL.p + (size_t)bidx * B * 3, sizeof(float) * B * 3,
                              hipMemcpyDeviceToDevice, M->stream));
        if ((rc = mlp_forward(M, B))) return rc;
        if ((rc = mlp_backward(M, B, cfg->w_att, cfg->w_mom))) return rc;
        M->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float lr = cfg->lr * (s < cfg->steps / 2 ? 1.0f : (s < 3 * cfg->steps / 4 ? 0.3f : 0.1f));
        k_adam<<<(NPARAM + 255) / 256, 256, 0, M->stream>>>(M->p.p, M->g.p, M->m1.p, M->m2.p, NPARAM, lr, b1, b2,
                                                            1e-8f, 1.0f - std::pow(b1, (float)M->step),
                                                            1.0f - std::pow(b2, (float)M->step));
        if ((rc = mlp_pack(M))) return rc;
        if (loss_hist && n_hist > 0 && (s % per == 0 || s == cfg->steps - 1)) {
            float lv = 0.0f;
            HIPCHK(hipMemcpyAsync(&lv, M->w.lossum.p, sizeof(float), hipMemcpyDeviceToHost, M->stream));
            HIPCHK(hipStreamSynchronize(M->stream));
            const int k = std::min(n_hist - 1, s / per);
            lsum[k] += lv / (float)B;
            lcnt[k] += 1;
        }
    }
    HIPCHK(hipStreamSynchronize(M->stream));
    if (loss_hist)
        for (int k = 0; k < n_hist; ++k) loss_hist[k] = lcnt[k] ? lsum[k] / (float)lcnt[k] : NAN;
    return SM2D_OK;
}

// Held-out accuracy: n samples drawn with `seed`; err3 = RMS error of (ln sigma_E, ln sigma_theta,
// attenuation exponent relative to the largest), err3[3..5] the largest absolute errors.
int sm2d_mlp_evaluate(sm2d_mlp* M, sm2d_ctx* c, int n, uint32_t seed, float* err6) {
    if (!M || !c || !err6 || n <= 0) return err("sm2d_mlp_evaluate: bad argument");
    if (!M->has_dict) return err("sm2d_mlp_evaluate: width dictionaries not set");
    HIPCHK(hipSetDevice(M->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int Bp = (n + 63) / 64 * 64;
    int rc = mlp_ensure(M, Bp);
    if (rc) return rc;
    DBuf<float> out3;
    HIPCHK(out3.alloc((size_t)Bp * 3));
    k_gen_samples<<<std::min((Bp + 255) / 256, 8192), 256, 0, M->stream>>>(c->tabs(), seed, Bp, M->w.X.p,
                                                                           M->w.labels.p, nullptr);
    if ((rc = mlp_forward(M, Bp))) return rc;
    k_decode<<<std::min((Bp + 255) / 256, 8192), 256, 0, M->stream>>>(M->w.Z4.p, Bp, M->dict, out3.p);
    HIPCHK(hipGetLastError());
    std::vector<float> pr((size_t)Bp * 3), lb((size_t)Bp * 3);
    HIPCHK(hipMemcpyAsync(pr.data(), out3.p, sizeof(float) * pr.size(), hipMemcpyDeviceToHost, M->stream));
    HIPCHK(hipMemcpyAsync(lb.data(), M->w.labels.p, sizeof(float) * lb.size(), hipMemcpyDeviceToHost, M->stream));
    HIPCHK(hipStreamSynchronize(M->stream));
    double s0 = 0, s1 = 0, s2 = 0, m0 = 0, m1 = 0, m2 = 0;
    for (int i = 0; i < n; ++i) {
        const double d0 = pr[3 * i] - lb[3 * i], d1 = pr[3 * i + 1] - lb[3 * i + 1];
        const double d2 = std::log((double)pr[3 * i + 2] + 1e-9) - (double)lb[3 * i + 2];
        s0 += d0 * d0;
        s1 += d1 * d1;
        s2 += d2 * d2;
        m0 = std::max(m0, std::fabs(d0));
        m1 = std::max(m1, std::fabs(d1));
        m2 = std::max(m2, std::fabs(d2));
    }
    err6[0] = (float)std::sqrt(s0 / n);
    err6[1] = (float)std::sqrt(s1 / n);
    err6[2] = (float)std::sqrt(s2 / n);
    err6[3] = (float)m0;
    err6[4] = (float)m1;
    err6[5] = (float)m2;
    return SM2D_OK;
}

// Surrogate table of a context: the network evaluated at every (material, energy bin, step node)
// the fine transport can query; materials are the distinct (rho, X0 ratio) rows of the context.
int sm2d_surrogate_attach(sm2d_ctx* c, sm2d_mlp* M, int n_nodes) {
    if (!c || !M || n_nodes < 2) return err("sm2d_surrogate_attach: bad argument");
    if (!c->tables) return err("sm2d_surrogate_attach: context tables not uploaded");
    if (!M->has_dict) return err("sm2d_surrogate_attach: the network has no width dictionaries");
    HIPCHK(hipSetDevice(c->device));
    const int Nz = c->g.Nz, NE = c->g.N_E;
    std::vector<float> rho(Nz), x0r(Nz);
    HIPCHK(hipMemcpyAsync(rho.data(), c->rho.p, sizeof(float) * Nz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(x0r.data(), c->x0r.p, sizeof(float) * Nz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<float> mr, mx;
    std::vector<int> rowmat(Nz, -1);
    for (int z = 0; z < Nz; ++z) {
        int m = -1;
        for (size_t k = 0; k < mr.size(); ++k)
            if (mr[k] == rho[z] && mx[k] == x0r[z]) m = (int)k;
        if (m < 0) {
            mr.push_back(rho[z]);
            mx.push_back(x0r[z]);
            m = (int)mr.size() - 1;
        }
        const bool in = rho[z] >= RHO_LO && rho[z] <= RHO_HI && x0r[z] >= X0R_LO && x0r[z] <= X0R_HI;
        rowmat[z] = in ? m : -1;
    }
    const int nmat = (int)mr.size();
    const int total = nmat * NE * n_nodes, Bp = (total + 63) / 64 * 64;
    int rc = mlp_ensure(M, Bp);
    if (rc) return rc;
    DBuf<float> dmr, dmx, out3;
    DBuf<uint8_t> valid;
    HIPCHK(dmr.alloc(nmat));
    HIPCHK(dmx.alloc(nmat));
    HIPCHK(out3.alloc((size_t)Bp * 3));
    HIPCHK(valid.alloc(Bp));
    HIPCHK(hipMemcpyAsync(dmr.p, mr.data(), sizeof(float) * nmat, hipMemcpyHostToDevice, M->stream));
    HIPCHK(hipMemcpyAsync(dmx.p, mx.data(), sizeof(float) * nmat, hipMemcpyHostToDevice, M->stream));
    HIPCHK(hipMemsetAsync(M->w.X.p, 0, sizeof(uint16_t) * (size_t)Bp * KP0, M->stream));
    k_table_inputs<<<std::min((total + 255) / 256, 8192), 256, 0, M->stream>>>(c->tabs(), dmr.p, dmx.p, nmat, n_nodes,
                                                                               M->w.X.p, valid.p);
    if ((rc = mlp_forward(M, Bp))) return rc;
    k_decode<<<std::min((Bp + 255) / 256, 8192), 256, 0, M->stream>>>(M->w.Z4.p, Bp, M->dict, out3.p);
    HIPCHK(hipGetLastError());
    std::vector<float> tab((size_t)Bp * 3);
    std::vector<uint8_t> hv(Bp);
    HIPCHK(hipMemcpyAsync(tab.data(), out3.p, sizeof(float) * tab.size(), hipMemcpyDeviceToHost, M->stream));
    HIPCHK(hipMemcpyAsync(hv.data(), valid.p, Bp, hipMemcpyDeviceToHost, M->stream));
    HIPCHK(hipStreamSynchronize(M->stream));
    for (int i = 0; i < total; ++i)
        if (!hv[i]) tab[3 * i] = tab[3 * i + 1] = tab[3 * i + 2] = NAN;
    tab.resize((size_t)total * 3);
    c->graphs_clear();
    HIPCHK(c->surr.alloc(tab.size()));
    HIPCHK(c->rowmat.alloc(Nz));
    HIPCHK(hipMemcpyAsync(c->surr.p, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->rowmat.p, rowmat.data(), sizeof(int) * Nz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->surr_nodes = n_nodes;
    c->surr_nmat = nmat;
    c->surr_on = 1;
    // the network itself, for the inline mode of sm2d_surrogate_enable (M must outlive that use)
    c->mlp_wb = M->wb.p;
    c->mlp_p = M->p.p;
    c->mlp_dict[0] = M->dict.lsE_lo;
    c->mlp_dict[1] = M->dict.lsE_hi;
    c->mlp_dict[2] = M->dict.lsT_lo;
    c->mlp_dict[3] = M->dict.lsT_hi;
    return surrogate_validate(c, c->surr_tol, nullptr);
}

int sm2d_surrogate_check(sm2d_ctx* c, float tol, float* stats4) {
    if (!c || !c->surr.p) return err("sm2d_surrogate_check: no surrogate table attached");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->surr_tol = tol;
    return surrogate_validate(c, tol, stats4);
}

int sm2d_surrogate_stats(sm2d_ctx* c, float* stats4) {
    if (!c || !stats4) return err("sm2d_surrogate_stats: null argument");
    for (int k = 0; k < 4; ++k) stats4[k] = c->surr_stats[k];
    return SM2D_OK;
}

int sm2d_surrogate_enable(sm2d_ctx* c, int on) {
    if (!c) return err("null context");
    if (on < 0 || on > 2) return err("sm2d_surrogate_enable: mode must be 0, 1 or 2");
    if (on && !c->surr.p) return err("sm2d_surrogate_enable: no surrogate table attached");
    if (on == 2 && !c->mlp_wb) return err("sm2d_surrogate_enable: no network attached (sm2d_surrogate_attach)");
    if (on != c->surr_on) c->graphs_clear();
    c->surr_on = on;
    return SM2D_OK;
}

// Exports the attached table (for the CPU oracle): table[n_mat][N_E][n_nodes][3] and the material
// index of every depth row (-1: outside the training range, plain physics).
int sm2d_surrogate_export(sm2d_ctx* c, float* table, int capacity, int* n_mat, int* n_nodes, int32_t* rowmat) {
    if (!c || !c->surr.p) return err("sm2d_surrogate_export: no surrogate table attached");
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)c->surr_nmat * c->g.N_E * c->surr_nodes * 3;
    if (n_mat) *n_mat = c->surr_nmat;
    if (n_nodes) *n_nodes = c->surr_nodes;
    if (table) {
        if ((size_t)capacity < n) return err("sm2d_surrogate_export: capacity");
        HIPCHK(hipMemcpy(table, c->surr.p, sizeof(float) * n, hipMemcpyDeviceToHost));
    }
    if (rowmat) HIPCHK(hipMemcpy(rowmat, c->rowmat.p, sizeof(int) * c->g.Nz, hipMemcpyDeviceToHost));
    return SM2D_OK;
}

// Installs a caller-built table in the layout sm2d_surrogate_export returns (for instance a
// table exported from one context and edited, or tabulated by other means).  Entries of a row
// whose material index is -1 are never read; NaN entries send the lookup back to the plain
// physics like the out-of-range rows.
int sm2d_surrogate_set_table(sm2d_ctx* c, const float* table, int n_mat, int n_nodes, const int32_t* rowmat) {
    if (!c || !table || !rowmat || n_mat < 1 || n_nodes < 2) return err("sm2d_surrogate_set_table: bad argument");
    if (!c->tables) return err("sm2d_surrogate_set_table: context tables not uploaded");
    for (int z = 0; z < c->g.Nz; ++z)
        if (rowmat[z] < -1 || rowmat[z] >= n_mat) return err("sm2d_surrogate_set_table: material index out of range");
    HIPCHK(hipSetDevice(c->device));
    const size_t n = (size_t)n_mat * c->g.N_E * n_nodes * 3;
    HIPCHK(hipStreamSynchronize(c->stream));
    c->graphs_clear();
    HIPCHK(c->surr.alloc(n));
    HIPCHK(c->rowmat.alloc(c->g.Nz));
    HIPCHK(hipMemcpyAsync(c->surr.p, table, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->rowmat.p, rowmat, sizeof(int) * c->g.Nz, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->surr_nodes = n_nodes;
    c->surr_nmat = n_mat;
    c->surr_on = 1;
    return SM2D_OK;
}

}  // extern "C"
