// tn_engine.hip -- host orchestration of the device kernels: the effective
// Hamiltonian operator, environment updates, Lanczos and the bond SVD.
#include "tn_engine.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

namespace tn {

void DBuf::ensure(size_t n)
{
    if (n <= bytes && p) return;
    // kernels queued on any stream may still read the old buffer
    if (p) {
        TN_HIP(hipDeviceSynchronize());
        TN_HIP(hipFree(p));
    }
    p = nullptr;
    bytes = 0;
    size_t want = n < 256 ? 256 : n;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        throw TnError{TN_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(want) +
                                        " bytes failed: " + hipGetErrorString(e)};
    }
    bytes = want;
}
void DBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

void sync(tn_ctx* c) { TN_HIP(hipStreamSynchronize(c->stream)); }
void h2d(tn_ctx* c, void* dst, const void* src, size_t bytes)
{
    TN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    sync(c);
}
void d2h_small(tn_ctx* c, double* dst, const double* src, int n)
{
    TN_REQUIRE(n >= 0 && n <= TN_PUBLISH_MAX && c->sc.h, TN_ERR_ARG, "d2h_small: bad size");
    const unsigned long long seq = ++c->sc.seq;
    TN_HIP(publish(src, n, c->sc.d, c->sc.flag_d, seq, c->stream));
    volatile unsigned long long* f = c->sc.flag_h;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned long spins = 1; *f != seq; spins++) {
        if ((spins & 0xffff) == 0) {
            // a long wait: hand over to the stream synchronisation, which also
            // reports a failed launch
            if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) {
                sync(c);
                TN_REQUIRE(*f == seq, TN_ERR_STATE, "d2h_small: completion flag not set");
                break;
            }
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    for (int i = 0; i < n; i++) dst[i] = ((volatile double*)c->sc.h)[i];
}

void d2h(tn_ctx* c, void* dst, const void* src, size_t bytes)
{
    TN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    sync(c);
}

// ---------------------------------------------------------------------------
// MPO mixing tables.  Orientation per op (W[w][w'][m][n], m bra, n ket):
//  HEFF : T1[a'][n][w'][b] -> T2[w][a'][m][b]      (sum over w', n)
//  ENVR : X [a'][n][w'][b] -> Y [a'][w][m][b]      (sum over w', n)
//  ENVL : X [a][w][n][b']  -> Y [a][m][w'][b']     (sum over w , n)
// ---------------------------------------------------------------------------
// GEMM dispatch with k-split and per-shape autotuning
void gemm_split(tn_ctx* c, const GemmDesc& g, int ks)
{
    if (ks <= 1) {
        TN_HIP(gemm(g, c->stream));
        return;
    }
    TN_REQUIRE(g.batch == 1 && g.K % (64 * ks) == 0, TN_ERR_ARG, "gemm_split: bad k-split");
    const int tacc = g.tin == T_F64 ? T_F64 : T_F32;
    const size_t es = tacc == T_F64 ? 8 : 4;
    c->splitk.ensure((size_t)ks * g.M * g.N * es);
    GemmDesc p = g;
    const int Kc = g.K / ks;
    p.K = Kc;
    p.batch = ks;
    p.sA = g.a_kmajor ? Kc : (long)Kc * g.lda;
    p.sB = g.b_kmajor ? Kc : (long)Kc * g.ldb;
    p.tout = tacc;
    p.C = c->splitk.p;
    p.ldc = g.N;
    p.sC = (long)g.M * g.N;
    p.accumulate = 0;
    TN_HIP(gemm(p, c->stream));
    TN_HIP(splitk_reduce(tacc, c->splitk.p, ks, g.M, g.N, g.tout, g.C, g.ldc, g.accumulate, c->stream));
}

static std::pair<int, int> tune_gemm(tn_ctx* c, const GemmDesc& g)
{
    hipStream_t s = c->stream;
    const int nt = gemm_num_tiles();
    const size_t es = dtype_size(g.tout);
    const size_t span = (size_t)(g.batch - 1) * g.sC + (size_t)(g.M - 1) * g.ldc + g.N;
    c->tune_C.ensure(span * es);
    GemmDesc t = g;
    t.C = c->tune_C.p;
    hipEvent_t e0, e1;
    TN_HIP(hipEventCreate(&e0));
    TN_HIP(hipEventCreate(&e1));
    std::pair<int, int> best{0, 1};
    float best_ms = 1e30f;
    // TN_TUNE_VERBOSE: every candidate's time on stderr
    static const bool verbose = getenv("TN_TUNE_VERBOSE") != nullptr;
    for (int code = 1; code <= nt; code++) {
        if (!gemm_tile_ok(code, g.tin, g.M, g.N, g.K)) continue;
        for (int ks : {1, 2, 4}) {
            if (ks > 1 && (g.batch != 1 || g.K % (64 * ks))) continue;
            t.tile = code;
            gemm_split(c, t, ks);  // warm
            TN_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < 3; r++) gemm_split(c, t, ks);
            TN_HIP(hipEventRecord(e1, s));
            TN_HIP(hipEventSynchronize(e1));
            float ms = 0;
            TN_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (verbose)
                fprintf(stderr, "tune tin %d tout %d M %d N %d K %d batch %d: tile %s ksplit %d %.4f ms\n", g.tin,
                        g.tout, g.M, g.N, g.K, g.batch, gemm_tile_name(code), ks, ms / 3);
            if (ms < best_ms) {
                best_ms = ms;
                best = {code, ks};
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return best;
}

std::pair<int, int> gemm_choice(tn_ctx* c, const GemmDesc& g)
{
    if (g.tile != 0) return {g.tile, 1};
    const std::array<long, 9> key{g.tin, g.tout, g.a_kmajor, g.b_kmajor, g.M, g.N, g.K, g.batch, g.accumulate};
    auto it = c->tune.find(key);
    if (it != c->tune.end()) return it->second;
    if (!c->autotune) return {gemm_default_tile(g.M, g.N, g.batch), 1};
    return c->tune.emplace(key, tune_gemm(c, g)).first->second;
}

void gemm_run(tn_ctx* c, const GemmDesc& g)
{
    if (g.tile != 0 || !c->autotune) {
        TN_HIP(gemm(g, c->stream));
        return;
    }
    const std::array<long, 9> key{g.tin, g.tout, g.a_kmajor, g.b_kmajor, g.M, g.N, g.K, g.batch, g.accumulate};
    auto it = c->tune.find(key);
    if (it == c->tune.end()) it = c->tune.emplace(key, tune_gemm(c, g)).first;
    GemmDesc t = g;
    t.tile = it->second.first;
    gemm_split(c, t, it->second.second);
}

void mix_build(DevMix& m, int op, const double* W, int dw1, int dw2, int p, int D1p, int D2p,
               hipStream_t s, int wl, int wr)
{
    if (op != MIX_HEFF && op != MIX_HEFF_T) wl = wr = -1;
    const size_t nw = (size_t)dw1 * dw2 * p * p;
    if (m.wsrc == W && m.op == op && m.D1p == D1p && m.D2p == D2p && m.p == p && m.dw1 == dw1 &&
        m.dw2 == dw2 && m.wl == wl && m.wr == wr && m.wkey.size() == nw &&
        std::memcmp(m.wkey.data(), W, nw * sizeof(double)) == 0)
        return;
    m.wkey.clear();  // set again once the new table is complete
    const bool ident = wl >= 0 || wr >= 0 || op == MIX_HEFF_T;
    const bool heff = op == MIX_HEFF || op == MIX_HEFF_T;
    // transposed stage-3 operand: row (m, b), column (w, a'), ldo = nL * D1p
    const long ldo = (long)(dw1 - (wl >= 0 ? 1 : 0)) * D1p;
    // right channels kept in the stage-1 output and left channels kept in the
    // stage-3 input, renumbered without the identity channel
    const int nR = dw2 - (wr >= 0 ? 1 : 0);
    auto cl = [&](int w) { return w - (wl >= 0 && w > wl ? 1 : 0); };
    auto cr = [&](int w) { return w - (wr >= 0 && w > wr ? 1 : 0); };
    std::vector<int> optr, esel, osel, ocode, ecode;
    std::vector<long> ooff, eoff;
    std::vector<double> ev;
    auto Wv = [&](int w, int wp, int mm, int n) { return W[(((long)w * dw2 + wp) * p + mm) * p + n]; };
    optr.push_back(0);
    if (heff || op == MIX_ENVR) {
        for (int w = 0; w < dw1; w++)
            for (int mm = 0; mm < p; mm++) {
                if (heff && w == wl) {
                    ooff.push_back((long)mm * D2p);  // y[a][mm][b]
                    osel.push_back(1);
                    ocode.push_back(mm | (1 << 16));
                } else {
                    if (op == MIX_HEFF_T) ooff.push_back((long)mm * D2p * ldo + (long)cl(w) * D1p);
                    else if (op == MIX_HEFF) ooff.push_back((long)cl(w) * D1p * p * D2p + (long)mm * D2p);
                    else ooff.push_back((long)w * p * D2p + (long)mm * D2p);
                    osel.push_back(0);
                    ocode.push_back(mm | (cl(w) << 8));
                }
                for (int wp = 0; wp < dw2; wp++)
                    for (int n = 0; n < p; n++) {
                        double v = Wv(w, wp, mm, n);
                        if (v != 0.0) {
                            if (wp == wr) {
                                eoff.push_back((long)n * D2p);  // x[a'][n][b]
                                esel.push_back(1);
                                ecode.push_back(n | (1 << 16));
                            } else {
                                eoff.push_back((long)n * nR * D2p + (long)cr(wp) * D2p);
                                esel.push_back(0);
                                ecode.push_back(n | (cr(wp) << 8));
                            }
                            ev.push_back(v);
                        }
                    }
                optr.push_back((int)ev.size());
            }
    } else {
        for (int wp = 0; wp < dw2; wp++)
            for (int mm = 0; mm < p; mm++) {
                ooff.push_back((long)mm * dw2 * D2p + (long)wp * D2p);
                for (int w = 0; w < dw1; w++)
                    for (int n = 0; n < p; n++) {
                        double v = Wv(w, wp, mm, n);
                        if (v != 0.0) {
                            eoff.push_back((long)w * p * D2p + (long)n * D2p);
                            ev.push_back(v);
                        }
                    }
                optr.push_back((int)ev.size());
            }
    }
    const int nout = (int)ooff.size(), nnz = (int)ev.size();
    std::vector<float> evf(ev.begin(), ev.end());
    // layout: optr | ooff | eoff | eval | evalf | esel | osel  (each 256-byte aligned)
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t o1 = 0, o2 = o1 + al((nout + 1) * sizeof(int)), o3 = o2 + al(nout * sizeof(long) + 8),
           o4 = o3 + al(nnz * sizeof(long) + 8), o5 = o4 + al(nnz * sizeof(double) + 8),
           o6 = o5 + al(nnz * sizeof(float) + 8), o7 = o6 + al(nnz * sizeof(int) + 8),
           o8 = o7 + al(nout * sizeof(int) + 8), o9 = o8 + al(nout * sizeof(int) + 8),
           tot = o9 + al(nnz * sizeof(int) + 8);
    const bool codes = op == MIX_HEFF_T;
    std::vector<char> host(tot, 0);
    std::memcpy(host.data() + o1, optr.data(), (nout + 1) * sizeof(int));
    if (nout) std::memcpy(host.data() + o2, ooff.data(), nout * sizeof(long));
    if (nnz) {
        std::memcpy(host.data() + o3, eoff.data(), nnz * sizeof(long));
        std::memcpy(host.data() + o4, ev.data(), nnz * sizeof(double));
        std::memcpy(host.data() + o5, evf.data(), nnz * sizeof(float));
    }
    if (ident) {
        if (nnz) std::memcpy(host.data() + o6, esel.data(), nnz * sizeof(int));
        if (nout) std::memcpy(host.data() + o7, osel.data(), nout * sizeof(int));
    }
    if (codes) {
        if (nout) std::memcpy(host.data() + o8, ocode.data(), nout * sizeof(int));
        if (nnz) std::memcpy(host.data() + o9, ecode.data(), nnz * sizeof(int));
    }
    m.buf.ensure(tot);
    TN_HIP(hipMemcpyAsync(m.buf.p, host.data(), tot, hipMemcpyHostToDevice, s));
    TN_HIP(hipStreamSynchronize(s));
    char* b = (char*)m.buf.p;
    m.t.nout = nout;
    m.t.nnz = nnz;
    m.t.optr = (const int*)(b + o1);
    m.t.ooff = (const long*)(b + o2);
    m.t.eoff = (const long*)(b + o3);
    m.t.eval = (const double*)(b + o4);
    m.t.evalf = (const float*)(b + o5);
    m.t.esel = ident ? (const int*)(b + o6) : nullptr;
    m.t.osel = ident ? (const int*)(b + o7) : nullptr;
    m.t.ocode = codes ? (const int*)(b + o8) : nullptr;
    m.t.ecode = codes ? (const int*)(b + o9) : nullptr;
    m.wsrc = W;
    m.wkey.assign(W, W + nw);
    m.op = op;
    m.D1p = D1p;
    m.D2p = D2p;
    m.p = p;
    m.dw1 = dw1;
    m.dw2 = dw2;
    m.wl = wl;
    m.wr = wr;
}

// ---------------------------------------------------------------------------
double HeffOp::flops_alg(int D1, int D2) const
{
    // 2 GEMM stages + W stage on logical dims: for D1 = D2 = D this is
    // 4 d Dw D^3 + 2 d^2 Dw^2 D^2 (one-site) with d -> p.
    double s1 = 2.0 * D1 * p * (double)Dw2 * D2 * D2;
    double s2 = 2.0 * (double)Dw1 * Dw2 * p * p * D1 * D2;
    double s3 = 2.0 * D1 * (double)p * D2 * Dw1 * D1;
    return s1 + s2 + s3;
}

double HeffOp::flops_gemm(int D1, int D2) const
{
    const int n2 = Dw2 - (wr_id >= 0 ? 1 : 0), n1 = Dw1 - (wl_id >= 0 ? 1 : 0);
    return 2.0 * D1 * p * (double)n2 * D2 * D2 + 2.0 * D1 * (double)p * D2 * n1 * D1;
}

static inline int lo_type(int prec) { return prec == TN_FP32 ? T_F32 : T_BF16; }

void gemm_run_fixed(tn_ctx* c, const GemmDesc& g)
{
    GemmDesc t = g;
    if (t.tile == 0) {
        const std::array<long, 9> key{g.tin, g.tout, g.a_kmajor, g.b_kmajor, g.M, g.N, g.K, g.batch, g.accumulate};
        auto it = c->tune.find(key);
        t.tile = it != c->tune.end() ? it->second.first : gemm_default_tile(g.M, g.N, g.batch);
    }
    TN_HIP(gemm(t, c->stream));
}

size_t heff_scratch_bytes(const HeffOp& op, size_t* t1, size_t* t2, size_t* xlo)
{
    const int nR = op.Dw2 - (op.wr_id >= 0 ? 1 : 0), nL = op.Dw1 - (op.wl_id >= 0 ? 1 : 0);
    const size_t n1 = (size_t)op.D1p * op.p * nR * op.D2p, n2 = (size_t)nL * op.D1p * op.p * op.D2p;
    const size_t es = op.precision == TN_FP64 ? 8 : lowprec_bytes(op.precision);
    const size_t a = op.precision == TN_FP64 ? n1 * 8 : n1 * 4;
    const size_t b = n2 * es;
    const size_t x = op.precision == TN_FP64 ? 0 : (size_t)op.n() * es;
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    if (t1) *t1 = up(a);
    if (t2) *t2 = up(b);
    if (xlo) *xlo = up(x);
    return up(a) + up(b) + up(x);
}

void heff_apply(tn_ctx* c, const HeffOp& op, const double* x, double* y, bool timed, const HeffScratch* scr)
{
    // a pending Lanczos scaling of x (x is then the driver's own basis vector,
    // scaled in place within the first pass over it), taken before anything
    // can throw
    const double pre = c->x_prescale;
    c->x_prescale = 0.0;
    double* const xs = pre != 0.0 ? const_cast<double*>(x) : nullptr;
    hipStream_t s = c->stream;
    const int p = op.p, D1p = op.D1p, D2p = op.D2p, Dw1 = op.Dw1, Dw2 = op.Dw2;
    const bool idl = op.wl_id >= 0, idr = op.wr_id >= 0;
    TN_REQUIRE(op.mix->op == MIX_HEFF_T && op.mix->wl == op.wl_id && op.mix->wr == op.wr_id, TN_ERR_STATE,
               "heff_apply: mixing table does not match the operator");
    TN_REQUIRE(!idl || (Dw1 >= 2 && (op.wl_id == 0 || op.wl_id == Dw1 - 1)), TN_ERR_STATE,
               "heff_apply: left identity channel must be the first or last of >= 2");
    TN_REQUIRE(!idr || (Dw2 >= 2 && (op.wr_id == 0 || op.wr_id == Dw2 - 1)), TN_ERR_STATE,
               "heff_apply: right identity channel must be the first or last of >= 2");
    // channels taken by the GEMMs: right [r0, r0 + nR), left [l0, l0 + nL)
    const int nR = Dw2 - (idr ? 1 : 0), r0 = op.wr_id == 0 ? 1 : 0;
    const int nL = Dw1 - (idl ? 1 : 0), l0 = op.wl_id == 0 ? 1 : 0;
    const long n = op.n();
    const long n1 = (long)D1p * p * nR * D2p;
    const long n2 = (long)nL * D1p * p * D2p;
    if (timed) TN_HIP(hipEventRecord(c->ev[0], s));
    GemmDesc g1{}, g3{};
    g1.a_kmajor = true;
    g1.b_kmajor = true;
    g1.M = D1p * p;
    g1.N = nR * D2p;
    g1.K = D2p;
    g1.batch = 1;
    g1.lda = D2p;
    g1.ldb = D2p;
    g1.ldc = (long)nR * D2p;
    g1.alpha = 1.0f;
    g1.tile = op.tile;
    // stage 3 as NT: the MPO stage writes T2 transposed, T2t[(m,b)][(w,a')]
    g3.a_kmajor = true;
    g3.b_kmajor = true;
    g3.M = D1p;
    g3.N = p * D2p;
    g3.K = nL * D1p;
    g3.batch = 1;
    g3.lda = (long)Dw1 * D1p;
    g3.ldb = (long)nL * D1p;
    g3.ldc = (long)p * D2p;
    g3.alpha = 1.0f;
    g3.tile = op.tile;
    g3.accumulate = idl ? 1 : 0;  // y already holds the identity channel's slice
    const long roff = (long)r0 * D2p * D2p;  // element offset of the first kept R channel
    const long loff = (long)l0 * D1p;        // column offset of the first kept L channel
    // work space: the caller's (allocation-free applies) or the context's
    void *T1 = nullptr, *T2 = nullptr, *XL = nullptr;
    if (scr) {
        size_t b1, b2, bx;
        heff_scratch_bytes(op, &b1, &b2, &bx);
        TN_REQUIRE(scr->T1 && scr->T2 && scr->T1_bytes >= b1 && scr->T2_bytes >= b2 &&
                       (bx == 0 || (scr->xlo && scr->xlo_bytes >= bx)),
                   TN_ERR_ARG, "heff_apply: work space too small");
        T1 = scr->T1;
        T2 = scr->T2;
        XL = scr->xlo;
    }
    auto run = [&](const GemmDesc& g) {
        if (scr) gemm_run_fixed(c, g);
        else gemm_run(c, g);
    };
    if (op.precision == TN_FP64) {
        if (!scr) {
            c->T1.ensure(n1 * 8);
            c->T2.ensure(n2 * 8);
            T1 = c->T1.p;
            T2 = c->T2.p;
        }
        if (xs) TN_HIP(scale_copy(pre, xs, xs, n, s));
        if (timed) TN_HIP(hipEventRecord(c->ev[1], s));
        g1.tin = T_F64; g1.tout = T_F64; g1.A = x; g1.B = op.R + roff; g1.C = T1;
        run(g1);
        if (timed) TN_HIP(hipEventRecord(c->ev[2], s));
        TN_HIP(wmix_idt(T_F64, T_F64, T_F64, T1, (long)p * nR * D2p, x, (long)p * D2p, T2,
                        (long)nL * D1p, y, (long)p * D2p, D1p, D2p, op.mix->t, s));
        if (timed) TN_HIP(hipEventRecord(c->ev[3], s));
        g3.tin = T_F64; g3.tout = T_F64; g3.A = op.L + loff; g3.B = T2; g3.C = y;
        run(g3);
    } else if (op.precision == TN_BF16X3) {
        // split-bfloat16 contractions: x, the environments and the MPO-stage
        // output as three bfloat16 planes, stage 1 into FP32, stage 3 into y
        // stage 1 with the MPO stage in its epilogue (no T1 in HBM) where the
        // fused kernel covers the shape; the allocation-free apply keeps the
        // two kernels (its work space is sized for them)
        static const bool no_fuse = getenv("TN_NO_FUSED_MIX") != nullptr;
        const bool fuse = !scr && !no_fuse && x3_stage1_mix_ok(p, nR, op.mix->t, D1p, D2p);
        if (!scr) {
            c->xlo.ensure((size_t)n * 6);
            if (!fuse) c->T1.ensure(n1 * 4);
            c->T2.ensure((size_t)n2 * 6);
            XL = c->xlo.p;
            T1 = c->T1.p;
            T2 = c->T2.p;
        }
        TN_REQUIRE(op.Lps > 0 && op.Rps > 0, TN_ERR_STATE, "heff_apply: split-bfloat16 operator without plane strides");
        if (xs) TN_HIP(split3_scale(xs, pre, XL, n, n, s));
        else TN_HIP(split3(T_F64, x, XL, n, n, s));
        if (timed) TN_HIP(hipEventRecord(c->ev[1], s));
        if (fuse) {
            TN_HIP(x3_stage1_mix(p, nR, (const bf16*)XL, n, (const bf16*)op.Rlo + roff, op.Rps, x, (bf16*)T2, n2,
                                 (long)nL * D1p, y, D1p, D2p, op.mix->t, s));
            if (timed) {
                TN_HIP(hipEventRecord(c->ev[2], s));
                TN_HIP(hipEventRecord(c->ev[3], s));
            }
        } else {
            g1.tin = T_BF16X3; g1.tout = T_F32; g1.A = XL; g1.psA = n;
            g1.B = (const bf16*)op.Rlo + roff; g1.psB = op.Rps; g1.C = T1;
            run(g1);
            if (timed) TN_HIP(hipEventRecord(c->ev[2], s));
            TN_HIP(wmix_idt(T_F32, T_F64, T_BF16X3, T1, (long)p * nR * D2p, x, (long)p * D2p, T2, (long)nL * D1p,
                            y, (long)p * D2p, D1p, D2p, op.mix->t, s, n2));
            if (timed) TN_HIP(hipEventRecord(c->ev[3], s));
        }
        g3.tin = T_BF16X3; g3.tout = T_F64; g3.A = (const bf16*)op.Llo + loff; g3.psA = op.Lps;
        g3.B = T2; g3.psB = n2; g3.C = y;
        run(g3);
    } else {
        const int tl = lo_type(op.precision);
        const size_t es = dtype_size(tl);
        if (!scr) {
            c->xlo.ensure(n * es);
            c->T1.ensure(n1 * 4);
            c->T2.ensure(n2 * es);
            XL = c->xlo.p;
            T1 = c->T1.p;
            T2 = c->T2.p;
        }
        if (xs) TN_HIP(convert_scale(xs, pre, tl, XL, n, s));
        else TN_HIP(convert(T_F64, x, tl, XL, n, s));
        if (timed) TN_HIP(hipEventRecord(c->ev[1], s));
        g1.tin = tl; g1.tout = T_F32; g1.A = XL; g1.B = (const char*)op.Rlo + roff * es; g1.C = T1;
        run(g1);
        if (timed) TN_HIP(hipEventRecord(c->ev[2], s));
        TN_HIP(wmix_idt(T_F32, tl, tl, T1, (long)p * nR * D2p, XL, (long)p * D2p, T2,
                        (long)nL * D1p, y, (long)p * D2p, D1p, D2p, op.mix->t, s));
        if (timed) TN_HIP(hipEventRecord(c->ev[3], s));
        g3.tin = tl; g3.tout = T_F64; g3.A = (const char*)op.Llo + loff * es; g3.B = T2; g3.C = y;
        run(g3);
    }
    if (timed) TN_HIP(hipEventRecord(c->ev[4], s));
}

void heff_timing_read(tn_ctx* c, double* gemm_ms, double* total_ms)
{
    float t01 = 0, t12 = 0, t23 = 0, t34 = 0;
    TN_HIP(hipEventSynchronize(c->ev[4]));
    TN_HIP(hipEventElapsedTime(&t01, c->ev[0], c->ev[1]));
    TN_HIP(hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
    TN_HIP(hipEventElapsedTime(&t23, c->ev[2], c->ev[3]));
    TN_HIP(hipEventElapsedTime(&t34, c->ev[3], c->ev[4]));
    if (gemm_ms) *gemm_ms += (double)t12 + (double)t34;
    if (total_ms) *total_ms += (double)t01 + t12 + t23 + t34;
}

void heff_stage_times(tn_ctx* c, double* s1, double* mix, double* s3)
{
    float t12 = 0, t23 = 0, t34 = 0;
    TN_HIP(hipEventSynchronize(c->ev[4]));
    TN_HIP(hipEventElapsedTime(&t12, c->ev[1], c->ev[2]));
    TN_HIP(hipEventElapsedTime(&t23, c->ev[2], c->ev[3]));
    TN_HIP(hipEventElapsedTime(&t34, c->ev[3], c->ev[4]));
    *s1 = t12;
    *mix = t23;
    *s3 = t34;
}

void make_lowprec(tn_ctx* c, int prec, const double* src, void* dst, long n, long ps)
{
    if (prec == TN_FP64) return;
    if (prec == TN_BF16X3) {
        TN_HIP(split3(T_F64, src, dst, n, ps > 0 ? ps : n, c->stream));
        return;
    }
    TN_HIP(convert(T_F64, src, lo_type(prec), dst, n, c->stream));
}

// ---------------------------------------------------------------------------
void env_left(tn_ctx* c, int prec, int p, int D1p, int D2p, int Dw1, int Dw2, const DevMix& mix,
              const double* Lin, const void* Lin_lo, const double* A, double* Lout, int tile)
{
    hipStream_t s = c->stream;
    const long nA = (long)D1p * p * D2p;
    const long nX = (long)D1p * Dw1 * p * D2p;
    const long nY = (long)D1p * p * Dw2 * D2p;
    GemmDesc e1{}, e3{};
    e1.a_kmajor = true; e1.b_kmajor = false;
    e1.M = D1p * Dw1; e1.N = p * D2p; e1.K = D1p; e1.batch = 1;
    e1.lda = D1p; e1.ldb = (long)p * D2p; e1.ldc = (long)p * D2p; e1.alpha = 1.0f; e1.tile = tile;
    e3.a_kmajor = false; e3.b_kmajor = false;
    e3.M = D2p; e3.N = Dw2 * D2p; e3.K = D1p * p; e3.batch = 1;
    e3.lda = D2p; e3.ldb = (long)Dw2 * D2p; e3.ldc = (long)Dw2 * D2p; e3.alpha = 1.0f; e3.tile = tile;
    if (prec == TN_FP64) {
        c->T1.ensure(nX * 8);
        c->T2.ensure(nY * 8);
        e1.tin = T_F64; e1.tout = T_F64; e1.A = Lin; e1.B = A; e1.C = c->T1.p;
        gemm_run(c, e1);
        TN_HIP(wmix(T_F64, T_F64, c->T1.p, (long)Dw1 * p * D2p, c->T2.p, (long)p * Dw2 * D2p, D1p,
                    D2p, mix.t, s));
        e3.tin = T_F64; e3.tout = T_F64; e3.A = A; e3.B = c->T2.p; e3.C = Lout;
        gemm_run(c, e3);
    } else {
        const int tl = lo_type(prec);
        const size_t es = dtype_size(tl);
        c->xlo.ensure(nA * es);
        c->T1.ensure(nX * 4);
        c->T2.ensure(nY * es);
        TN_HIP(convert(T_F64, A, tl, c->xlo.p, nA, s));
        e1.tin = tl; e1.tout = T_F32; e1.A = Lin_lo; e1.B = c->xlo.p; e1.C = c->T1.p;
        gemm_run(c, e1);
        TN_HIP(wmix(T_F32, tl, c->T1.p, (long)Dw1 * p * D2p, c->T2.p, (long)p * Dw2 * D2p, D1p,
                    D2p, mix.t, s));
        e3.tin = tl; e3.tout = T_F64; e3.A = c->xlo.p; e3.B = c->T2.p; e3.C = Lout;
        gemm_run(c, e3);
    }
}

void env_right(tn_ctx* c, int prec, int p, int D1p, int D2p, int Dw1, int Dw2, const DevMix& mix,
               const double* Rin, const void* Rin_lo, const double* A, double* Rout, int tile)
{
    hipStream_t s = c->stream;
    const long nA = (long)D1p * p * D2p;
    const long nX = (long)D1p * p * Dw2 * D2p;
    const long nY = (long)D1p * Dw1 * p * D2p;
    GemmDesc f1{}, f3{};
    f1.a_kmajor = true; f1.b_kmajor = true;
    f1.M = D1p * p; f1.N = Dw2 * D2p; f1.K = D2p; f1.batch = 1;
    f1.lda = D2p; f1.ldb = D2p; f1.ldc = (long)Dw2 * D2p; f1.alpha = 1.0f; f1.tile = tile;
    f3.a_kmajor = true; f3.b_kmajor = true;
    f3.M = D1p; f3.N = D1p; f3.K = p * D2p; f3.batch = Dw1;
    f3.lda = (long)p * D2p; f3.sA = 0;
    f3.ldb = (long)Dw1 * p * D2p; f3.sB = (long)p * D2p;
    f3.ldc = D1p; f3.sC = (long)D1p * D1p;
    f3.alpha = 1.0f; f3.tile = tile;
    if (prec == TN_FP64) {
        c->T1.ensure(nX * 8);
        c->T2.ensure(nY * 8);
        f1.tin = T_F64; f1.tout = T_F64; f1.A = A; f1.B = Rin; f1.C = c->T1.p;
        gemm_run(c, f1);
        TN_HIP(wmix(T_F64, T_F64, c->T1.p, (long)p * Dw2 * D2p, c->T2.p, (long)Dw1 * p * D2p, D1p,
                    D2p, mix.t, s));
        f3.tin = T_F64; f3.tout = T_F64; f3.A = A; f3.B = c->T2.p; f3.C = Rout;
        gemm_run(c, f3);
    } else {
        const int tl = lo_type(prec);
        const size_t es = dtype_size(tl);
        c->xlo.ensure(nA * es);
        c->T1.ensure(nX * 4);
        c->T2.ensure(nY * es);
        TN_HIP(convert(T_F64, A, tl, c->xlo.p, nA, s));
        f1.tin = tl; f1.tout = T_F32; f1.A = c->xlo.p; f1.B = Rin_lo; f1.C = c->T1.p;
        gemm_run(c, f1);
        TN_HIP(wmix(T_F32, tl, c->T1.p, (long)p * Dw2 * D2p, c->T2.p, (long)Dw1 * p * D2p, D1p,
                    D2p, mix.t, s));
        f3.tin = tl; f3.tout = T_F64; f3.A = c->xlo.p; f3.B = c->T2.p; f3.C = Rout;
        gemm_run(c, f3);
    }
}

// ---------------------------------------------------------------------------
// Host tridiagonal Ritz solve (k <= a few hundred): the two lowest eigenvalues
// of T = tridiag(b, a, b) by Sturm-sequence bisection, the eigenvector of the
// lowest by inverse iteration with a partially pivoted LU of T - t0 I.
namespace {
int sturm_count(int k, const double* a, const double* b2, double x)
{
    int cnt = 0;
    double q = a[0] - x;
    if (q < 0) cnt++;
    for (int i = 1; i < k; i++) {
        if (q == 0.0) q = 1e-300;
        q = a[i] - x - b2[i - 1] / q;
        if (q < 0) cnt++;
    }
    return cnt;
}
}  // namespace

// Lowest eigenvalue of tridiag(b, a, b) by Laguerre's iteration from below
// (the Gershgorin bound): for a polynomial with real roots it increases
// monotonically to the smallest root and converges cubically.  G = f'/f and
// H = -(ln f)'' come from the LDL^T pivots d_i of T - x, which stay positive
// below the smallest root.  Returns false (and leaves x) if it does not settle.
static bool laguerre_lowest(int k, const double* a, const double* b2, double lo, double span, double* x)
{
    double xv = lo;
    for (int it = 0; it < 60; it++) {
        double d = a[0] - xv, dp = -1.0, dpp = 0.0;
        if (!(d > 0.0)) break;
        double G = dp / d, H = (dp / d) * (dp / d) - dpp / d;
        for (int i = 1; i < k; i++) {
            const double q = b2[i - 1] / d;
            const double dn = a[i] - xv - q;
            const double dpn = -1.0 + q * dp / d;
            const double dppn = q * (dpp * d - 2.0 * dp * dp) / (d * d);
            d = dn;
            dp = dpn;
            dpp = dppn;
            if (!(d > 0.0)) break;
            G += dp / d;
            H += (dp / d) * (dp / d) - dpp / d;
        }
        if (!(d > 0.0)) {
            // at (rounding) the root: a pivot reached zero
            *x = xv;
            return true;
        }
        const double n = (double)k;
        const double disc = std::max(0.0, (n - 1.0) * (n * H - G * G));
        const double den = G - std::sqrt(disc);  // G < 0 below every root
        if (!(den < 0.0)) break;
        const double step = -n / den;
        xv += step;
        if (step <= 4e-16 * (std::fabs(xv) + 1e-3 * span)) {
            *x = xv;
            return true;
        }
    }
    return false;
}

void tridiag_lowest(int k, const double* a, const double* b, double* t0, double* t1, double* z, bool want_t1)
{
    thread_local std::vector<double> work;
    thread_local std::vector<unsigned char> piv;
    work.resize(8 * (size_t)std::max(k, 1));
    piv.resize(std::max(k, 1));
    double* b2 = work.data();
    double *dd = b2 + k, *du = dd + k, *dl = du + k, *du2 = dl + k, *rhs = du2 + k, *zz = rhs + k;
    for (int i = 0; i + 1 < k; i++) b2[i] = b[i] * b[i];
    double lo = a[0], hi = a[0];
    for (int i = 0; i < k; i++) {
        const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i < k - 1 ? std::fabs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    const double span = hi - lo + 1e-300;
    lo -= 1e-10 * span + 1e-300;
    hi += 1e-10 * span + 1e-300;
    auto eig = [&](int m) {
        double l = lo, h = hi;
        for (int it = 0; it < 200; it++) {
            const double mid = 0.5 * (l + h);
            if (mid <= l || mid >= h) break;
            if (sturm_count(k, a, b2, mid) > m) h = mid;
            else l = mid;
        }
        return 0.5 * (l + h);
    };
    double x0;
    if (!laguerre_lowest(k, a, b2, lo, span, &x0)) x0 = eig(0);
    *t0 = x0;
    if (t1) *t1 = (want_t1 && k > 1) ? eig(1) : *t0;
    if (k == 1) {
        z[0] = 1.0;
        return;
    }
    double scale = 0.0;
    for (int i = 0; i < k; i++) scale = std::max(scale, std::fabs(a[i]) + (i < k - 1 ? std::fabs(b[i]) : 0.0));
    const double pert = (scale > 0 ? scale : 1.0) * 1e-14;
    for (int i = 0; i < k; i++) zz[i] = 1.0 / std::sqrt((double)k) * (1.0 + 0.01 * ((i * 7919) % 13) / 13.0);
    for (int iter = 0; iter < 3; iter++) {
        for (int i = 0; i < k; i++) {
            dd[i] = a[i] - *t0;
            du2[i] = 0.0;
            if (i < k - 1) du[i] = dl[i] = b[i];
        }
        for (int i = 0; i < k - 1; i++) {
            if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
                piv[i] = 0;
                if (dd[i] == 0.0) dd[i] = pert;
                const double f = dl[i] / dd[i];
                dl[i] = f;
                dd[i + 1] -= f * du[i];
            } else {
                piv[i] = 1;
                const double f = dd[i] / dl[i];
                dd[i] = dl[i];
                dl[i] = f;
                const double tmp = du[i];
                du[i] = dd[i + 1];
                dd[i + 1] = tmp - f * dd[i + 1];
                if (i < k - 2) {
                    du2[i] = du[i + 1];
                    du[i + 1] = -f * du[i + 1];
                }
            }
        }
        if (dd[k - 1] == 0.0) dd[k - 1] = pert;
        for (int i = 0; i < k; i++) rhs[i] = zz[i];
        for (int i = 0; i < k - 1; i++) {
            if (piv[i] == 0) {
                rhs[i + 1] -= dl[i] * rhs[i];
            } else {
                const double t = rhs[i];
                rhs[i] = rhs[i + 1];
                rhs[i + 1] = t - dl[i] * rhs[i + 1];
            }
        }
        rhs[k - 1] /= dd[k - 1];
        rhs[k - 2] = (rhs[k - 2] - du[k - 2] * rhs[k - 1]) / dd[k - 2];
        for (int i = k - 3; i >= 0; i--) rhs[i] = (rhs[i] - du[i] * rhs[i + 1] - du2[i] * rhs[i + 2]) / dd[i];
        double nr = 0.0;
        for (int i = 0; i < k; i++) nr += rhs[i] * rhs[i];
        nr = std::sqrt(nr);
        for (int i = 0; i < k; i++) zz[i] = rhs[i] / nr;
    }
    for (int i = 0; i < k; i++) z[i] = zz[i];
}

// ---------------------------------------------------------------------------
// Lanczos with full re-orthogonalisation: the three-term step removes the
// large components of H v_j (alpha, beta), then one classical Gram-Schmidt
// pass against every Krylov vector removes the rounding residue, which is
// small, so a single pass keeps the basis orthonormal to working precision
// (a single pass on the raw H v_j would not: its error grows by |H v_j|/beta
// per step).  Krylov vectors, Ritz vector and residual stay in FP64 on the
// device; per iteration the host reads the coefficients and the norm (one
// sync) and solves the tridiagonal problem itself.
//
// Partial re-orthogonalisation (H. D. Simon, Math. Comp. 42 (1984) 115-142):
// each step removes the components along v_j and v_{j-1} only, and the host
// propagates estimates omega_{j+1,k} of v_{j+1}.v_k through the recurrence
//   beta_j w'_k = beta_k w_{k+1} + (alpha_k - alpha_j) w_k + beta_{k-1} w_{k-1}
//                 - beta_{j-1} omega_{j-1,k}  (+ rounding, eps_op |T|),
// with w = omega_{j,.}.  When an estimate exceeds sqrt(eps_op) the new vector
// gets a Gram-Schmidt pass against the whole basis, and so does the next one
// (the recurrence couples consecutive vectors).  Semi-orthogonality at
// sqrt(eps_op) keeps the tridiagonal matrix the projection of the operator to
// O(eps_op |H|), the operator's own accuracy.
LanczosResult lanczos(tn_ctx* c, ApplyFn fn, const void* op, long N, double* x, double tol,
                      int max_krylov, int max_restarts, double tol_rel, bool partial, double eps_op, bool prescale)
{
    hipStream_t s = c->stream;
    LanczosResult res;
    const int kmax = std::max(2, std::min(max_krylov, 500));
    c->krylov.ensure((size_t)(kmax + 1) * N * 8);
    double* V = c->krylov.as<double>();
    c->lanczos_small.ensure(sizeof(double) * (4 * 512 + 64));
    double* h = c->lanczos_small.as<double>();  // [0, 512): coefficients, [512, 1024): Ritz coefficients
    double* zd = h + 512;
    double* nrm = h + 1024;
    const long pcap = mdot_part_size(N, kmax + 1);
    c->part.ensure(pcap * 8);
    double* part = c->part.as<double>();
    std::vector<double> al(kmax), be(kmax), z(kmax), hh(kmax + 4);
    // fused step kernels: 16-byte vector access needs an even length
    const bool fused = N % 2 == 0 && kmax <= 512;
    const bool pro = partial && fused;
    const double eta = std::sqrt(eps_op);
    std::vector<double> om_old(kmax + 2, 0.0), om(kmax + 2, 0.0), om_new(kmax + 2, 0.0);
    // the scaling of the next basis vector not launched yet (prescale)
    double pending = 0.0;
    c->x_prescale = 0.0;

    auto read = [&](double* dst, const double* src, int n) {
        if (n <= TN_PUBLISH_MAX) d2h_small(c, dst, src, n);
        else d2h(c, dst, src, (size_t)n * 8);
    };
    auto norm_host = [&](const double* v) {
        TN_HIP(mdot(v, N, 1, v, N, nrm, part, (int)pcap, s));
        double v2 = 0;
        read(&v2, nrm, 1);
        return std::sqrt(v2);
    };

    for (int restart = 0; restart <= max_restarts; restart++) {
        double nx = norm_host(x);
        if (!(nx > 0.0) || !std::isfinite(nx)) {
            // zero or non-finite start: seeded pseudo-random start vector
            std::vector<double> r(N);
            uint64_t st = 0x5eedULL + restart;
            for (long i = 0; i < N; i++) {
                st += 0x9e3779b97f4a7c15ULL;
                uint64_t zz = st;
                zz = (zz ^ (zz >> 30)) * 0xbf58476d1ce4e5b9ULL;
                zz = (zz ^ (zz >> 27)) * 0x94d049bb133111ebULL;
                zz ^= zz >> 31;
                r[i] = (double)(zz >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
            }
            h2d(c, x, r.data(), N * 8);
            nx = norm_host(x);
        }
        TN_HIP(scale_copy(1.0 / nx, x, V, N, s));
        int k = 0;
        bool conv = false;
        bool force = false;     // the previous vector had a full pass: this one gets one too
        double tnorm = 0.0;     // Gershgorin bound on |T|
        om[0] = 1.0;
        for (int j = 0; j < kmax; j++) {
            double* vj = V + (long)j * N;
            double* w = V + (long)(j + 1) * N;
            if (pending != 0.0) {  // v_j = w_{j-1} / beta_{j-1}, inside the apply
                c->x_prescale = pending;
                pending = 0.0;
            }
            try {
                fn(c, op, vj, w);
            } catch (...) {
                c->x_prescale = 0.0;
                throw;
            }
            res.n_matvec++;
            // three-term step with the exact alpha = v_j.w (h[0] on the device),
            // then one classical Gram-Schmidt pass against every Krylov vector
            // (h[1..j+1]) removes what rounding left; h[j+2] = |w|^2
            TN_HIP(mdot(vj, N, 1, w, N, h, part, (int)pcap, s));
            if (pro) {
                // the same step and pass against v_{j-1}, v_j only
                const int kk = j >= 1 ? 2 : 1;
                const double* Vl = V + (long)(j + 1 - kk) * N;
                TN_HIP(lz_step_dots(w, Vl, N, kk, h, j > 0 ? be[j - 1] : 0.0, N, h + 1, part, pcap, s));
                TN_HIP(lz_update_norm(w, Vl, N, kk, h + 1, N, h + kk + 1, part, pcap, s));
                read(hh.data(), h, kk + 2);
                al[j] = hh[0] + hh[kk];
                be[j] = std::sqrt(std::max(hh[kk + 1], 0.0));
                tnorm = std::max(tnorm, std::fabs(al[j]) + be[j] + (j > 0 ? be[j - 1] : 0.0));
                const double noise = eps_op * tnorm;
                double wmax = 0.0;
                if (be[j] > 0.0) {
                    for (int q = 0; q + 1 < j; q++) {
                        double t = be[q] * om[q + 1] + (al[q] - al[j]) * om[q] - be[j - 1] * om_old[q];
                        if (q > 0) t += be[q - 1] * om[q - 1];
                        t = (t + std::copysign(noise, t)) / be[j];
                        om_new[q] = t;
                        wmax = std::max(wmax, std::fabs(t));
                    }
                    for (int q = std::max(j - 1, 0); q <= j; q++) om_new[q] = noise / be[j];
                }
                om_new[j + 1] = 1.0;
                if (j > 0 && (force || wmax > eta || !std::isfinite(wmax))) {
                    TN_HIP(mdot(V, N, j + 1, w, N, h + 1, part, (int)pcap, s));
                    TN_HIP(lz_update_norm(w, V, N, j + 1, h + 1, N, h + j + 2, part, pcap, s));
                    read(hh.data(), h, j + 3);
                    al[j] += hh[j + 1];
                    const double bc = std::sqrt(std::max(hh[j + 2], 0.0));
                    // only rounding left: an invariant subspace (breakdown)
                    be[j] = bc < 1e-10 * be[j] ? 0.0 : bc;
                    for (int q = 0; q <= j; q++) om_new[q] = be[j] > 0.0 ? noise / be[j] : 0.0;
                    force = !force;
                    res.n_full++;
                } else {
                    force = false;
                }
                std::swap(om_old, om);
                std::swap(om, om_new);
            } else if (fused) {
                // three-term step fused with the Gram-Schmidt dot products, then
                // the Gram-Schmidt update fused with the norm
                TN_HIP(lz_step_dots(w, V, N, j + 1, h, j > 0 ? be[j - 1] : 0.0, N, h + 1, part, pcap, s));
                TN_HIP(lz_update_norm(w, V, N, j + 1, h + 1, N, h + j + 2, part, pcap, s));
            } else {
                TN_HIP(three_term(w, vj, j > 0 ? V + (long)(j - 1) * N : nullptr, h, j > 0 ? be[j - 1] : 0.0, N,
                                  s));
                TN_HIP(mdot(V, N, j + 1, w, N, h + 1, part, (int)pcap, s));
                TN_HIP(gemv_acc(V, N, j + 1, h + 1, -1.0, 1.0, w, N, s));
                TN_HIP(mdot(w, N, 1, w, N, h + j + 2, part, (int)pcap, s));
            }
            if (!pro) {
                read(hh.data(), h, j + 3);
                al[j] = hh[0] + hh[j + 1];
                be[j] = std::sqrt(std::max(hh[j + 2], 0.0));
                tnorm = std::max(tnorm, std::fabs(al[j]) + be[j] + (j > 0 ? be[j - 1] : 0.0));
                res.n_full++;
            }
            tridiag_lowest(j + 1, al.data(), be.data(), &res.e0, nullptr, z.data(), false);
            k = j + 1;
            res.residual = std::fabs(be[j] * z[j]);
            if (!std::isfinite(res.e0)) throw TnError{TN_ERR_STATE, "Lanczos: non-finite Ritz value"};
            // breakdown: the Krylov space is invariant up to rounding, judged
            // against the operator's scale |T| (the remainder of a complete
            // space is ~ eps |T| sqrt(N), not ~ eps |theta|)
            const bool done = res.residual < tol || res.residual < tol_rel * std::fabs(res.e0) ||
                              be[j] < 1e-14 * std::max(1.0, std::fabs(res.e0)) || be[j] <= 1e-13 * tnorm;
            if (done || j + 1 == kmax)  // the second Ritz value is only reported
                tridiag_lowest(j + 1, al.data(), be.data(), &res.e0, &res.e1, z.data(), true);
            if (done) {
                conv = true;
                break;
            }
            if (j + 1 < kmax) {
                if (prescale) pending = 1.0 / be[j];
                else TN_HIP(scale_copy(1.0 / be[j], w, w, N, s));
            }
        }
        // Ritz vector x = sum_i z_i v_i, normalised
        h2d(c, zd, z.data(), (size_t)k * 8);
        TN_HIP(gemv_acc(V, N, k, zd, 1.0, 0.0, x, N, s));
        const double nr = norm_host(x);
        TN_HIP(scale_copy(1.0 / nr, x, x, N, s));
        if (conv) {
            res.converged = true;
            break;
        }
    }
    return res;
}

// ---------------------------------------------------------------------------
// Householder QR of the column-major pp x qp matrix A (ld pp; pp >= qp, both
// multiples of 32), blocked by 32 columns: each panel is factored by one
// workgroup (hqr_panel), then the block reflector I - V T V^T is applied to
// the trailing columns by three FP64 GEMMs,
//   A_t <- A_t - V (T^T (V^T A_t)),
// over rows c0.. only (V is zero above).  R overwrites A (zeros below the
// diagonal).  Q (pp x qp, ld pp) is formed backwards from [I; 0]:
//   Q[c0:, c0:] <- Q[c0:, c0:] - V (T (V^T Q[c0:, c0:])).
// H holds the reflectors (pp x qp), T the panels' 32 x 32 factors, W1 and W2
// qp x 32 each.
static void house_qr(tn_ctx* c, double* A, int pp, int qp, double* Q, double* H, double* T, double* W1,
                     double* W2)
{
    hipStream_t s = c->stream;
    TN_REQUIRE(pp % 32 == 0 && qp % 32 == 0 && pp >= qp && qp > 0, TN_ERR_ARG, "house_qr: bad shape");
    // Y (nt columns, ld pp, rows c0..pp-1) <- (I - V F V^T) Y, F = T^T for the
    // factorisation (H_31 ... H_0 applied) or T for the Q formation
    // (H_0 ... H_31), computed transposed: Y^T -= ((Y^T V) F^T) V^T
    auto apply = [&](double* Y, int nt, int c0, const double* Tk, bool qform) {
        const double* Vk = H + (long)c0 * pp + c0;
        GemmDesc g{};
        g.tin = T_F64; g.tout = T_F64; g.batch = 1; g.alpha = 1.0f;
        // W1 = Y^T V  (nt x 32, row-major)
        g.a_kmajor = true; g.b_kmajor = true;
        g.M = nt; g.N = 32; g.K = pp - c0;
        g.A = Y; g.lda = pp; g.B = Vk; g.ldb = pp; g.C = W1; g.ldc = 32;
        TN_HIP(gemm(g, s));
        // W2 = W1 T (the reflector's transpose on A) or W1 T^T (Q formation)
        g.a_kmajor = true; g.b_kmajor = qform;
        g.M = nt; g.N = 32; g.K = 32;
        g.A = W1; g.lda = 32; g.B = Tk; g.ldb = 32; g.C = W2; g.ldc = 32;
        TN_HIP(gemm(g, s));
        // Y^T -= W2 V^T
        g.a_kmajor = true; g.b_kmajor = false;
        g.M = nt; g.N = pp - c0; g.K = 32;
        g.A = W2; g.lda = 32; g.B = Vk; g.ldb = pp; g.C = Y; g.ldc = pp;
        g.alpha = -1.0f; g.accumulate = 1;
        TN_HIP(gemm(g, s));
    };
    const int np = qp / 32;
    for (int k = 0; k < np; k++) {
        const int c0 = 32 * k;
        TN_HIP(hqr_panel(A, pp, pp, c0, H + (long)c0 * pp, pp, T + (long)k * 1024, s));
        if (qp - c0 - 32 > 0) apply(A + (long)(c0 + 32) * pp + c0, qp - c0 - 32, c0, T + (long)k * 1024, false);
    }
    TN_HIP(fill_zero(Q, (long)pp * qp * 8, s));
    TN_HIP(set_identity(Q, pp, qp, s));
    for (int k = np - 1; k >= 0; k--) {
        const int c0 = 32 * k;
        apply(Q + (long)c0 * pp + c0, qp - c0, c0, T + (long)k * 1024, true);
    }
}

// ---------------------------------------------------------------------------
// Thin SVD by one-sided block Jacobi on the q <= p columns of X0 (p x q):
// X0 = A (r >= c) or A^T (r < c).
//
// Small q: X = X0, X J = Q S, so X0 = Q S J^T.
//
// q >= 128: QR-preconditioned (Z. Drmac, K. Veselic, SIAM J. Matrix Anal.
// Appl. 29 (2008) 1322-1342): X0 P = Q_h R by Householder QR of the columns
// sorted by norm (P), then the Jacobi runs on X = P R^T, X J = W S, so
// X0 = (Q_h J) S W^T.  With a fast-decaying spectrum (every two-site split)
// the plain iteration needs 20-50 sweeps; on P R^T it needs about 10.
// largest cosine between two columns below which the plain iteration runs
static const double kSvdPlainCos = 0.1;

int svd_device(tn_ctx* c, const double* A, long rs, long cs, int r, int cc, double* U, long ldu,
               int zr, double* Vt, long ldvt, int zk, std::vector<double>& s_out)
{
    hipStream_t s = c->stream;
    const int k = std::min(r, cc);
    const bool tall = r >= cc;
    const int p = tall ? r : cc, q = tall ? cc : r;
    const int qp = pad32(q);
    static const bool plain_only = getenv("TN_SVD_PLAIN") != nullptr;
    const bool big = q >= 128 && !plain_only;
    const int pp = pad32(p);
    c->svd_V.ensure((size_t)qp * qp * 8);
    double* V = c->svd_V.as<double>();
    c->svd_s.ensure((size_t)(qp + 64) * 8 * 2 + 1024 * 8);
    double* amax_d = c->svd_s.as<double>() + 2 * qp + 128;
    double* nd = c->svd_s.as<double>();
    c->part.ensure(1024 * 8);
    double amax = 0.0;
    // X0 (column j of X0 at B + j*ldb), scaled to max |entry| = 1 (Gram
    // entries of tiny matrices would underflow)
    auto load_scaled = [&](double* B, long ldb, int rows_out) {
        if (tall) TN_HIP(copy2(A, cs, rs, q, p, B, ldb, 1, qp, rows_out, s));
        else TN_HIP(copy2(A, rs, cs, q, p, B, ldb, 1, qp, rows_out, s));
        TN_HIP(absmax(B, ldb * qp, amax_d, c->part.as<double>(), 1024, s));
        d2h(c, &amax, amax_d, 8);
        if (!std::isfinite(amax)) throw TnError{TN_ERR_STATE, "svd: non-finite matrix entries"};
        if (amax > 0.0) TN_HIP(scale_by_inv(B, amax_d, ldb * qp, s));
    };
    // the Jacobi iterate X (pj rows, column j at X + j*ldx)
    double* X = nullptr;
    long ldx = p;
    int pj = p;
    bool pre = false;
    double* Qh = nullptr;  // pre: Q_h (pp x qp, ld pp)
    if (big) {
        c->svd_R.ensure((size_t)pp * qp * 8);
        c->svd_Q.ensure((size_t)pp * qp * 8);
        c->svd_H.ensure((size_t)pp * qp * 8);
        double* R = c->svd_R.as<double>();
        double* Hb = c->svd_H.as<double>();
        load_scaled(Hb, pp, pp);
        // precondition only columns far from orthogonal: the largest cosine
        // between two columns, from the Gram matrix (one GEMM).  Nearly
        // orthogonal columns (a one-site gauge move's site tensor) converge
        // in a few plain sweeps, where the QR would cost more than it saves.
        GemmDesc g{};
        g.tin = T_F64; g.tout = T_F64; g.batch = 1; g.alpha = 1.0f;
        g.a_kmajor = true; g.b_kmajor = true;
        g.M = qp; g.N = qp; g.K = pp;
        g.A = Hb; g.lda = pp; g.B = Hb; g.ldb = pp; g.C = R; g.ldc = qp;
        TN_HIP(gemm(g, s));
        TN_HIP(max_cos(R, qp, q, nd, s));
        std::vector<double> mc(q);
        d2h(c, mc.data(), nd, (size_t)q * 8);
        const double cmax = *std::max_element(mc.begin(), mc.end());
        pre = !(cmax < kSvdPlainCos);
        if (getenv("TN_SVD_TRACE")) fprintf(stderr, "# svdcos q=%d cmax=%.3e pre=%d\n", q, cmax, (int)pre);
        if (pre) {
            c->svd_X.ensure((size_t)qp * qp * 8);
            X = c->svd_X.as<double>();
            ldx = qp;
            pj = qp;
            const size_t wt = (size_t)2 * qp * 32 + (size_t)(qp / 32) * 1024;
            c->svd_W.ensure((wt + (size_t)qp / 2 + 8) * 8);
            Qh = c->svd_Q.as<double>();
            double* W = c->svd_W.as<double>();
            int* dpiv = (int*)(W + wt);
            // static column pivoting: the QR sees the columns by decreasing
            // norm (the grading the preconditioning relies on; a two-site
            // split's columns come in blocks by physical index, and its
            // unsorted QR gave the Jacobi no fewer sweeps than the plain
            // iteration)
            TN_HIP(col_norms(Hb, pp, p, q, nd, s));
            std::vector<double> cn(q);
            d2h(c, cn.data(), nd, (size_t)q * 8);
            std::vector<int> piv(qp);
            std::iota(piv.begin(), piv.end(), 0);
            std::stable_sort(piv.begin(), piv.begin() + q, [&](int a, int b) { return cn[a] > cn[b]; });
            h2d(c, dpiv, piv.data(), (size_t)qp * sizeof(int));
            TN_HIP(gather_cols(Hb, pp, R, pp, dpiv, pp, qp, s));
            house_qr(c, R, pp, qp, Qh, Hb, W + (size_t)2 * qp * 32, W, W + (size_t)qp * 32);
            // X0 P = Q_h R, so X0 = Q_h (P R^T)^T: X = P R^T,
            // X[j*qp + piv[i]] = R(j, i) = R[i*pp + j]
            TN_HIP(tr_scatter(R, pp, X, qp, dpiv, qp, s));
        } else {
            // the plain iteration on the loaded copy (rows p..pp-1 are zero)
            X = Hb;
            ldx = pp;
        }
    } else {
        c->svd_X.ensure((size_t)p * qp * 8);
        X = c->svd_X.as<double>();
        load_scaled(X, ldx, p);
    }
    TN_HIP(set_identity(V, qp, qp, s));
    // tournament pairs for nb = qp/16 column blocks
    const int nb = qp / 16;
    std::vector<int> pairs((size_t)(nb - 1) * nb);
    for (int rd = 0; rd < nb - 1; rd++)
        for (int i = 0; i < nb / 2; i++) {
            auto elem = [&](int kp) { return kp == 0 ? 0 : ((kp - 1 + rd) % (nb - 1)) + 1; };
            pairs[(size_t)rd * nb + 2 * i] = elem(i);
            pairs[(size_t)rd * nb + 2 * i + 1] = elem(nb - 1 - i);
        }
    c->svd_pairs.ensure(pairs.size() * sizeof(int) + 64);
    h2d(c, c->svd_pairs.p, pairs.data(), pairs.size() * sizeof(int));
    c->svd_flag.ensure(256);
    int* flag = c->svd_flag.as<int>();
    const double tol = std::max(1e-15, 4.4e-16 * std::sqrt((double)pj));
    int hflag = 1;
    // one sweep of rotations on each pair's Gram matrix per visit: the pairs
    // come back every outer sweep, and a round has only nb/2 pairs, so its
    // latency (mostly the inner rounds) sets the cost.  Measured on the
    // 512 x 512 two-site splits at D=256: one inner sweep halves the SVD time
    // against diagonalising each pair (up to 20 inner sweeps), with more
    // outer sweeps and the same singular values.
    const int inner = 1;
    // workgroups per pair, each rotating its own rows (about 256 in all, at
    // least 32 rows each); the iterate alternates between X and X2
    const int nchunk = std::max(1, std::min({16, 256 / std::max(1, nb / 2), std::max(1, pj / 32)}));
    c->svd_X2.ensure((size_t)ldx * qp * 8);
    double* Xa = X;
    double* Xb = c->svd_X2.as<double>();
    for (int sweep = 0; sweep < 80 && hflag; sweep++) {
        TN_HIP(set_int(flag, 0, s));
        for (int rd = 0; rd < nb - 1; rd++) {
            TN_HIP(jacobi_step(Xa, Xb, ldx, pj, V, qp, qp, c->svd_pairs.as<int>() + (size_t)rd * nb, nb / 2,
                               nchunk, tol, flag, s, inner));
            std::swap(Xa, Xb);
        }
        TN_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, s));
        sync(c);
    }
    X = Xa;
    // singular values = column norms; sort descending
    double* sd = c->svd_s.as<double>();
    TN_HIP(col_norms(X, ldx, pj, q, sd, s));
    std::vector<double> sv(q);
    TN_HIP(hipMemcpyAsync(sv.data(), sd, q * 8, hipMemcpyDeviceToHost, s));
    sync(c);
    std::vector<int> perm(q);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return sv[a] > sv[b]; });
    const double smax = q > 0 ? sv[perm[0]] : 0.0;
    std::vector<double> inv(qp, 1.0);
    std::vector<char> good(q, 0);
    for (int j = 0; j < q; j++) {
        if (sv[j] > 1e-300 && sv[j] > 1e-140 * smax) {
            good[j] = 1;
            inv[j] = 1.0 / sv[j];
        }
    }
    // normalise the good columns of X
    double* dinv = sd + qp + 32;
    h2d(c, dinv, inv.data(), qp * 8);
    TN_HIP(scale_rows(X, ldx, q, pj, dinv, s));
    // orthonormal completion of (exactly) zero columns, in the sorted order,
    // within the first pc rows (pre: the right factor lives in q dimensions;
    // its left factor Q_h J is orthonormal in every column already)
    const int pc = pre ? q : p;
    for (int jj = 0; jj < q; jj++) {
        const int j = perm[jj];
        if (good[j]) continue;
        if (pc <= 0) break;
        double* xj = X + (long)j * ldx;
        long pcap = mdot_part_size(pc, q);
        c->part.ensure(pcap * 8);
        c->pack.ensure((size_t)2 * q * 8 + 256);
        double* h = c->pack.as<double>();
        double* hc = h + q + 16;
        for (int tries = 0; tries < pc && !good[j]; tries++) {
            TN_HIP(set_unit_col(xj, pc, (jj + tries) % pc, s));
            for (int pass = 0; pass < 2; pass++) {
                std::vector<double> coef(q);
                TN_HIP(mdot(X, ldx, q, xj, pc, h, c->part.as<double>(), (int)pcap, s));
                TN_HIP(hipMemcpyAsync(coef.data(), h, q * 8, hipMemcpyDeviceToHost, s));
                sync(c);
                for (int t = 0; t < q; t++)
                    if (!good[t] || t == j) coef[t] = 0.0;
                h2d(c, hc, coef.data(), q * 8);
                TN_HIP(gemv_acc(X, ldx, q, hc, -1.0, 1.0, xj, pc, s));
            }
            double nr2 = 0;
            TN_HIP(mdot(xj, pc, 1, xj, pc, h, c->part.as<double>(), (int)pcap, s));
            TN_HIP(hipMemcpyAsync(&nr2, h, 8, hipMemcpyDeviceToHost, s));
            sync(c);
            if (nr2 > 0.25) {
                double f = 1.0 / std::sqrt(nr2);
                std::vector<double> one(1, f);
                h2d(c, hc, one.data(), 8);
                TN_HIP(scale_rows(xj, pc, 1, pc, hc, s));
                sync(c);
                good[j] = 1;
            }
        }
    }
    // gather into U / Vt (zero padded first)
    if (zr > 0) TN_HIP(fill_zero(U, (long)zr * ldu * 8, s));
    if (zk > 0) TN_HIP(fill_zero(Vt, (long)zk * ldvt * 8, s));
    int* dperm = (int*)(sd + 2 * qp + 64);
    c->svd_perm.ensure((size_t)(q + 64) * 12);
    dperm = c->svd_perm.as<int>();
    double* ones = (double*)(dperm + ((q + 63) / 64) * 64);
    std::vector<double> onev(k, 1.0);
    h2d(c, dperm, perm.data(), k * sizeof(int));
    h2d(c, ones, onev.data(), k * 8);
    // the factor with p-long columns (L, ld ldl) and the one with q-long
    // columns (Rt, ld qp): plain X and J, preconditioned Q_h J and X
    const double* Lf = X;
    long ldl = ldx;
    const double* Rf = V;
    if (pre) {
        double* QJ = c->svd_H.as<double>();
        GemmDesc g{};
        g.tin = T_F64; g.tout = T_F64; g.batch = 1; g.alpha = 1.0f;
        // J has drifted from orthogonality by the rounding of its rotations
        // (~1e-13 after ten sweeps at q = 512), which Q_h J would inherit: one
        // Newton-Schulz step J <- J + J (I - J^T J) / 2 squares the drift away.
        // E = I - J^T J in the R buffer, the polished J through QJ's buffer.
        double* E = c->svd_R.as<double>();
        TN_HIP(fill_zero(E, (long)qp * qp * 8, s));
        TN_HIP(set_identity(E, qp, qp, s));
        g.a_kmajor = true; g.b_kmajor = true;
        g.M = qp; g.N = qp; g.K = qp;
        g.A = V; g.lda = qp; g.B = V; g.ldb = qp; g.C = E; g.ldc = qp;
        g.alpha = -1.0f; g.accumulate = 1;
        TN_HIP(gemm(g, s));
        // (J + J E / 2)^T = J^T + E J^T / 2 (E symmetric), row-major
        TN_HIP(hipMemcpyAsync(QJ, V, (size_t)qp * qp * 8, hipMemcpyDeviceToDevice, s));
        g.a_kmajor = true; g.b_kmajor = false;
        g.A = E; g.lda = qp; g.B = V; g.ldb = qp; g.C = QJ; g.ldc = qp;
        g.alpha = 0.5f; g.accumulate = 1;
        TN_HIP(gemm(g, s));
        TN_HIP(hipMemcpyAsync(V, QJ, (size_t)qp * qp * 8, hipMemcpyDeviceToDevice, s));
        g.alpha = 1.0f; g.accumulate = 0;
        // Q_h J into the reflector buffer: (Q_h J)^T = J^T Q_h^T, row-major
        g.a_kmajor = true; g.b_kmajor = false;
        g.M = qp; g.N = pp; g.K = qp;
        g.A = V; g.lda = qp; g.B = Qh; g.ldb = pp; g.C = QJ; g.ldc = pp;
        TN_HIP(gemm(g, s));
        Lf = QJ;
        ldl = pp;
        Rf = X;
    }
    if (tall) {
        // U (r x k) from the p-side factor, Vt (k x c) from the q-side one
        TN_HIP(svd_gather(Lf, ldl, p, Rf, qp, q, dperm, ones, k, U, ldu, 1, Vt, ldvt, 1, s));
    } else {
        // Vt (k x c) from the p-side factor (columns of length c), U (r x k) from the q-side one
        TN_HIP(svd_gather(Lf, ldl, p, Rf, qp, q, dperm, ones, k, Vt, 1, ldvt, U, 1, ldu, s));
    }
    s_out.resize(k);
    for (int j = 0; j < k; j++) s_out[j] = sv[perm[j]] * (amax > 0.0 ? amax : 1.0);
    sync(c);
    return k;
}

// ---------------------------------------------------------------------------
// CholeskyQR (G = A^T A, G = L L^T, Q = A L^-T) with every large step a GEMM
// on the matrix cores.  The 64x64 diagonal blocks of L are factored and
// inverted in LDS (chol_diag); L^-1 is completed by recursive doubling of
// block pairs, inv([L11 0; L21 L22]) = [L11^-1 0; -L22^-1 L21 L11^-1 L22^-1],
// one batched GEMM pair per level, so Q is a single GEMM.

// G = A^T A (n x n, row-major, ld n)
static void qr_gram(tn_ctx* c, const double* A, long lda, int m, int n, double* G)
{
    GemmDesc g{};
    g.tin = T_F64; g.tout = T_F64; g.a_kmajor = false; g.b_kmajor = false;
    g.M = n; g.N = n; g.K = m; g.batch = 1;
    g.A = A; g.lda = lda; g.B = A; g.ldb = lda; g.C = G; g.ldc = n; g.alpha = 1.0f;
    gemm_run(c, g);
}

// G (destroyed) = L L^T; L and Linv lower triangular n x n.  A pivot <= 0
// sets flag[0].  shift > 0 adds shift * trace(G) to the logical diagonal first.
static void qr_factor(tn_ctx* c, double* G, int n, int nl, double shift, double* L, double* Linv, int* flag)
{
    hipStream_t s = c->stream;
    if (shift > 0.0) TN_HIP(shift_diag(G, n, nl, shift, s));
    TN_HIP(add_diag(G, n, nl, n, 1.0, s));  // identity on the zero padding
    TN_HIP(fill_zero(L, (long)n * n * 8, s));
    TN_HIP(fill_zero(Linv, (long)n * n * 8, s));
    for (int k0 = 0; k0 < n; k0 += 64) {
        const int nbk = std::min(64, n - k0), r0 = k0 + nbk, rem = n - r0;
        TN_HIP(chol_diag(G, n, k0, nbk, L, n, Linv, n, flag, s));
        if (rem <= 0) continue;
        GemmDesc p{};  // L[r0:, k0:r0] = G[r0:, k0:r0] Linv_kk^T
        p.tin = T_F64; p.tout = T_F64; p.a_kmajor = true; p.b_kmajor = true;
        p.M = rem; p.N = nbk; p.K = nbk; p.batch = 1;
        p.A = G + (size_t)r0 * n + k0; p.lda = n;
        p.B = Linv + (size_t)k0 * n + k0; p.ldb = n;
        p.C = L + (size_t)r0 * n + k0; p.ldc = n; p.alpha = 1.0f;
        gemm_run(c, p);
        GemmDesc t{};  // G[r0:, r0:] -= Lp Lp^T
        t.tin = T_F64; t.tout = T_F64; t.a_kmajor = true; t.b_kmajor = true;
        t.M = rem; t.N = rem; t.K = nbk; t.batch = 1;
        t.A = L + (size_t)r0 * n + k0; t.lda = n;
        t.B = L + (size_t)r0 * n + k0; t.ldb = n;
        t.C = G + (size_t)r0 * n + r0; t.ldc = n; t.alpha = -1.0f; t.accumulate = 1;
        gemm_run(c, t);
    }
    // recursive doubling: at the start of level b every aligned b-block (and
    // the shorter tail block) of Linv holds the inverse of that block of L
    double* X = c->qr_T.as<double>();
    for (int b = 64; b < n; b *= 2) {
        const int P = n / (2 * b), tail = n - P * 2 * b;
        for (int part = 0; part < 2; part++) {
            const int np = part == 0 ? P : (tail > b ? 1 : 0);
            if (np == 0) continue;
            const int s2 = part == 0 ? b : tail - b;  // size of L22
            const size_t o = part == 0 ? 0 : (size_t)P * 2 * b;
            GemmDesc x{};  // X = L21 L11^-1   (s2 x b)
            x.tin = T_F64; x.tout = T_F64; x.a_kmajor = true; x.b_kmajor = false;
            x.M = s2; x.N = b; x.K = b; x.batch = np;
            x.A = L + (o + b) * n + o; x.lda = n; x.sA = 2L * b * n + 2L * b;
            x.B = Linv + o * n + o; x.ldb = n; x.sB = 2L * b * n + 2L * b;
            x.C = X; x.ldc = b; x.sC = (long)b * b; x.alpha = 1.0f;
            gemm_run(c, x);
            GemmDesc y{};  // Linv21 = -L22^-1 X
            y.tin = T_F64; y.tout = T_F64; y.a_kmajor = true; y.b_kmajor = false;
            y.M = s2; y.N = b; y.K = s2; y.batch = np;
            y.A = Linv + (o + b) * n + o + b; y.lda = n; y.sA = 2L * b * n + 2L * b;
            y.B = X; y.ldb = b; y.sB = (long)b * b;
            y.C = Linv + (o + b) * n + o; y.ldc = n; y.sC = 2L * b * n + 2L * b; y.alpha = -1.0f;
            gemm_run(c, y);
        }
    }
}

// Q = A Linv^T  (m x n)
static void qr_apply(tn_ctx* c, const double* A, long lda, int m, int n, const double* Linv, double* Q)
{
    GemmDesc q{};
    q.tin = T_F64; q.tout = T_F64; q.a_kmajor = true; q.b_kmajor = true;
    q.M = m; q.N = n; q.K = n; q.batch = 1;
    q.A = A; q.lda = lda; q.B = Linv; q.ldb = n; q.C = Q; q.ldc = n; q.alpha = 1.0f;
    gemm_run(c, q);
}

// Lacc <- Lacc Lk  (A = Q_prev Lacc^T = Q Lk^T Lacc^T)
static void qr_accumulate(tn_ctx* c, double* Lacc, const double* Lk, int n)
{
    GemmDesc g{};
    g.tin = T_F64; g.tout = T_F64; g.a_kmajor = true; g.b_kmajor = false;
    g.M = n; g.N = n; g.K = n; g.batch = 1;
    g.A = Lacc; g.lda = n; g.B = Lk; g.ldb = n; g.C = c->qr_T.p; g.ldc = n; g.alpha = 1.0f;
    gemm_run(c, g);
    TN_HIP(hipMemcpyAsync(Lacc, c->qr_T.p, (size_t)n * n * 8, hipMemcpyDeviceToDevice, c->stream));
}

// Second pass without a factorisation: with G = Q1^T Q1 = I + E and
// |E|_F <= 1e-3, Q = Q1 G^{-1/2} and A = Q1 L1^T = Q (L1 G^{1/2})^T, the
// inverse square root and the square root from their binomial series, taken
// to the lowest order k whose remainder c_{k+1} |E|^{k+1} (c = 3/8, 5/16,
// 35/128 for the inverse square root, the larger of the two) stays below
// 3e-13, so no Gram check follows: order 1 for |E| <= 8.9e-7, 2 for
// |E| <= 9.9e-5, 3 up to 1e-3.  With |E|_F <= 1e-13 Q1 is kept as it is.  The
// factor is no longer triangular; a gauge move does not need it to be.
// Returns false (nothing written) when |E| is too large for the series.
static bool polar_finish(tn_ctx* c, const double* Q1, int m, int n, int nl, double* Q, double* Lacc)
{
    hipStream_t s = c->stream;
    double* G = c->qr_G.as<double>();
    c->qr_W.ensure((size_t)n * n * 8);
    double* E = c->qr_T.as<double>();
    double* E2 = c->qr_L.as<double>();
    double* E3 = c->qr_linv.as<double>();
    double* Xm = G;
    double* Ym = c->qr_W.as<double>();
    qr_gram(c, Q1, n, m, n, G);
    TN_HIP(eye_sub(G, n, nl, E, s));
    const long nn = (long)n * n;
    const long pcap = mdot_part_size(nn, 1);
    c->part.ensure(std::max<long>(pcap, 1024) * 8);
    double* h = c->svd_s.as<double>();
    TN_HIP(mdot(E, nn, 1, E, nn, h, c->part.as<double>(), (int)pcap, s));
    double e2 = 0;
    d2h_small(c, &e2, h, 1);
    const double e = std::sqrt(e2);
    if (!(e <= 1e-3)) return false;
    if (e <= 1e-13) {
        if (Q != Q1) TN_HIP(hipMemcpyAsync(Q, Q1, (size_t)m * n * 8, hipMemcpyDeviceToDevice, s));
        return true;
    }
    const int order = e <= 8.9e-7 ? 1 : (e <= 9.9e-5 ? 2 : 3);
    GemmDesc g{};
    g.tin = T_F64; g.tout = T_F64; g.a_kmajor = true; g.b_kmajor = false;
    g.M = n; g.N = n; g.K = n; g.batch = 1; g.lda = n; g.ldb = n; g.ldc = n; g.alpha = 1.0f;
    if (order >= 2) {
        g.A = E; g.B = E; g.C = E2;
        gemm_run(c, g);
    }
    if (order >= 3) {
        g.A = E2; g.B = E; g.C = E3;
        gemm_run(c, g);
    }
    // unused powers are read with a zero coefficient
    const double ci[3] = {-0.5, order >= 2 ? 0.375 : 0.0, order >= 3 ? -0.3125 : 0.0};
    const double cs[3] = {0.5, order >= 2 ? -0.125 : 0.0, order >= 3 ? 0.0625 : 0.0};
    TN_HIP(series3(E, order >= 2 ? E2 : E, order >= 3 ? E3 : E, n, ci, cs, Xm, Ym, s));
    GemmDesc q{};
    q.tin = T_F64; q.tout = T_F64; q.a_kmajor = true; q.b_kmajor = false;
    q.M = m; q.N = n; q.K = n; q.batch = 1;
    q.A = Q1; q.lda = n; q.B = Xm; q.ldb = n; q.C = Q; q.ldc = n; q.alpha = 1.0f;
    gemm_run(c, q);
    qr_accumulate(c, Lacc, Ym, n);
    return true;
}

// Shifted CholeskyQR: every pass factors X^T X directly and, if a pivot
// breaks down (cond(X) beyond ~1e8), again with the shift
// 11 (m n + n (n+1)) u ||X||_F^2 (Fukaya, Kannan, Nakatsukasa, Yamamoto,
// Yanagisawa, SIAM J. Sci. Comput. 42 (2020) A477), which keeps cond of the
// pass's output moderate.  At least two passes run; from the third on the Gram
// matrix of the current Q is checked first and the result accepted when
// max|Q^T Q - I| <= tol.  At most five passes.
bool cholqr_device(tn_ctx* c, const double* A, long lda, int m, int n, int nl, double* Q, double* Lacc,
                   double tol, int* shifted, bool triangular)
{
    hipStream_t s = c->stream;
    TN_REQUIRE(m % 32 == 0 && n % 32 == 0 && m >= n && nl <= n, TN_ERR_ARG, "cholqr: bad shape");
    c->qr_G.ensure((size_t)n * n * 8);
    c->qr_L.ensure((size_t)n * n * 8);
    c->qr_T.ensure((size_t)n * n * 8);
    c->qr_linv.ensure((size_t)n * n * 8);
    c->qr_Q.ensure((size_t)m * n * 8);
    c->svd_flag.ensure(256);
    c->part.ensure(1024 * 8);
    c->svd_s.ensure(64 * 8);
    double* G = c->qr_G.as<double>();
    double* Lk = c->qr_L.as<double>();
    double* Linv = c->qr_linv.as<double>();
    double* Q1 = c->qr_Q.as<double>();
    int* flag = c->svd_flag.as<int>();
    double* dev = c->svd_s.as<double>();
    const double shift = 11.0 * ((double)m * n + (double)n * (n + 1)) * std::ldexp(1.0, -53);
    const int max_pass = 5;
    if (shifted) *shifted = 0;
    const double* X = A;
    long ldx = lda;
    for (int pass = 1;; pass++) {
        qr_gram(c, X, ldx, m, n, G);
        if (pass >= 3) {
            TN_HIP(eye_deviation(G, n, nl, dev, c->part.as<double>(), s));
            double h = 0;
            d2h_small(c, &h, dev, 1);
            if (std::isfinite(h) && h <= tol) {
                if (X != Q) TN_HIP(hipMemcpyAsync(Q, X, (size_t)m * n * 8, hipMemcpyDeviceToDevice, s));
                return true;
            }
            if (pass > max_pass) return false;
        }
        double* Lp = pass == 1 ? Lacc : Lk;
        TN_HIP(set_int(flag, 0, s));
        qr_factor(c, G, n, nl, 0.0, Lp, Linv, flag);
        int hflag = 0;
        d2h(c, &hflag, flag, sizeof(int));
        if (hflag) {
            TN_HIP(set_int(flag, 0, s));
            qr_gram(c, X, ldx, m, n, G);
            qr_factor(c, G, n, nl, shift, Lp, Linv, flag);
            d2h(c, &hflag, flag, sizeof(int));
            if (hflag) return false;
            if (shifted) *shifted = 1;
        }
        double* Y = X == Q1 ? Q : Q1;
        qr_apply(c, X, ldx, m, n, Linv, Y);
        if (pass > 1) qr_accumulate(c, Lacc, Lk, n);
        X = Y;
        ldx = n;
        if (pass == 1 && !triangular && polar_finish(c, Q1, m, n, nl, Q, Lacc)) return true;
    }
}

}  // namespace tn
