// tn_batch.hip -- many independent chains swept in lockstep.
//
// Round t of a pass gives every chain that still has sites its t-th site
// update (eigensolve, gauge move, environment update).  Every stage of a round
// is one grouped launch over the active chains:
//   * the Heff stages and the environment updates are persistent grouped GEMMs
//     (one launch per tile shape; the tile of each problem comes from the
//     per-shape tuning table, i.e. per (D1, D2, d, Dw) bucket);
//   * the Lanczos vector steps (three-term step fused with the Gram-Schmidt
//     dots, Gram-Schmidt update fused with the norm, scaling) are batched
//     kernels with one job per chain, and one polled read back per Lanczos
//     iteration serves every chain;
//   * an environment channel equal to the identity is applied as a copy, as
//     in the single-chain operator (one batched check per round);
//   * the CholeskyQR gauge move runs its 64x64 diagonal blocks for all chains
//     in one launch (one workgroup per chain), so their latency is paid once
//     per block step instead of once per chain.
// Contractions are FP32 on the matrix cores against FP32 environment copies;
// Krylov vectors, Ritz solve, gauge and energies stay FP64.  A chain whose
// CholeskyQR misses the orthogonality bound takes the single-chain gauge path
// (shifted passes, SVD fallback) for that site.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <thread>

#include "tn_chain.h"

using namespace tn;

struct tn_batch {
    tn_ctx* ctx = nullptr;
    std::vector<tn_chain*> ch;
    int kmax = 0;
    size_t stride = 0;  // doubles of Lanczos coefficients per chain
    // per chain scratch
    std::vector<DBuf> kry, T1, T2, xlo, qX, qQ1, qQ, qG, qL, qLk, qLinv, qT, qLacc;
    DBuf coef, part, flags;
    // descriptor ring: pinned host staging, device copy; reused from the
    // start only right after a host synchronisation
    char* ring_h = nullptr;
    DBuf ring_d;
    size_t ring_cap = 0, ring_off = 0;
    size_t stage_begin = 0;  // start of the staged, not yet copied, descriptors
    std::vector<double> hcoef;
    // host threads for the per-chain work between launches
    struct Pool* pool = nullptr;
};

// A few host threads for the per-chain host work of a lockstep iteration
// (tridiagonal Ritz solves of every active chain).  run(n, f) calls f(i) for
// i < n on the workers and the calling thread and returns when all are done;
// the first exception is rethrown on the caller.
struct Pool {
    std::vector<std::thread> th;
    std::mutex m;
    std::condition_variable cv, done;
    std::function<void(int)> fn;
    std::atomic<int> next{0};
    int n = 0, busy = 0;
    unsigned long long gen = 0;
    bool stop = false;
    std::exception_ptr err;

    explicit Pool(int workers)
    {
        for (int t = 0; t < workers; t++) th.emplace_back([this] { loop(); });
    }
    ~Pool()
    {
        {
            std::lock_guard<std::mutex> lk(m);
            stop = true;
        }
        cv.notify_all();
        for (auto& t : th) t.join();
    }
    void work()
    {
        for (int i; (i = next.fetch_add(1)) < n;) {
            try {
                fn(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(m);
                if (!err) err = std::current_exception();
            }
        }
    }
    void loop()
    {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return stop || gen != seen; });
                if (stop) return;
                seen = gen;
            }
            work();
            std::lock_guard<std::mutex> lk(m);
            if (--busy == 0) done.notify_one();
        }
    }
    void run(int items, const std::function<void(int)>& f)
    {
        if (items < 16 || th.empty()) {
            for (int i = 0; i < items; i++) f(i);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m);
            fn = f;
            n = items;
            next = 0;
            busy = (int)th.size();
            err = nullptr;
            gen++;
        }
        cv.notify_all();
        work();
        std::unique_lock<std::mutex> lk(m);
        done.wait(lk, [&] { return busy == 0; });
        if (err) std::rethrow_exception(err);
    }
};

namespace {

// copy the staged descriptors [stage_begin, ring_off) to the device, as one
// transfer ordered before the launches that follow on the stream
void flush(tn_batch* b)
{
    if (b->ring_off > b->stage_begin)
        TN_HIP(hipMemcpyAsync((char*)b->ring_d.p + b->stage_begin, b->ring_h + b->stage_begin,
                              b->ring_off - b->stage_begin, hipMemcpyHostToDevice, b->ctx->stream));
    b->stage_begin = b->ring_off;
}

// the ring may be reused from its start: the caller has just synchronised
// with the stream, after every staged list was flushed
void ring_reset(tn_batch* b)
{
    TN_REQUIRE(b->stage_begin == b->ring_off, TN_ERR_STATE, "batch: descriptors staged but not copied");
    b->ring_off = b->stage_begin = 0;
}

void batch_sync(tn_batch* b)
{
    flush(b);
    sync(b->ctx);
    b->ring_off = b->stage_begin = 0;
}

// place a descriptor list in the ring; the device copy is valid after the next
// flush (stage several lists, flush once, then launch)
template <typename T>
const T* stage(tn_batch* b, const std::vector<T>& v)
{
    const size_t bytes = v.size() * sizeof(T);
    const size_t span = (bytes + 255) & ~(size_t)255;
    TN_REQUIRE(span <= b->ring_cap, TN_ERR_ARG, "batch: descriptor list larger than the ring");
    if (b->ring_off + span > b->ring_cap) batch_sync(b);
    char* h = b->ring_h + b->ring_off;
    std::memcpy(h, v.data(), bytes);
    const T* d = (const T*)((char*)b->ring_d.p + b->ring_off);
    b->ring_off += span;
    return d;
}

template <typename T>
const T* push(tn_batch* b, const std::vector<T>& v)
{
    const T* d = stage(b, v);
    flush(b);
    return d;
}

// read n doubles back; waits for the stream to reach the read (polled
// completion flag up to TN_PUBLISH_MAX values), which also frees the ring
void readback(tn_batch* b, double* dst, const double* src, size_t n)
{
    flush(b);
    if (n <= (size_t)TN_PUBLISH_MAX) d2h_small(b->ctx, dst, src, (int)n);
    else d2h(b->ctx, dst, src, n * 8);
    b->ring_off = b->stage_begin = 0;
}

// A list of GEMMs sharing types and layouts, as one grouped launch per tile:
// gemm_group_stage places the descriptors, gemm_group_launch launches them
// (after a flush).
struct GroupLaunch {
    int tin, tout;
    bool akm, bkm;
    int tile, bk;
    const GemmProb* probs;
    const int* off;
    int nprob, total;
};

std::vector<GroupLaunch> gemm_group_stage(tn_batch* b, const std::vector<GemmDesc>& list)
{
    std::vector<GroupLaunch> out;
    if (list.empty()) return out;
    tn_ctx* c = b->ctx;
    std::map<std::pair<int, int>, std::vector<int>> groups;
    for (int q = 0; q < (int)list.size(); q++) {
        const GemmDesc& g = list[q];
        TN_REQUIRE(g.batch == 1 && g.tin == list[0].tin && g.tout == list[0].tout &&
                       g.a_kmajor == list[0].a_kmajor && g.b_kmajor == list[0].b_kmajor,
                   TN_ERR_ARG, "gemm_group: mixed problem types");
        int tile = gemm_choice(c, g).first;
        if (!gemm_tile_ok(tile, g.tin, g.M, g.N, g.K)) tile = gemm_default_tile(g.M, g.N, 1);
        groups[{tile, gemm_k_step(g.tin, g.K)}].push_back(q);
    }
    for (const auto& kv : groups) {
        const int tile = kv.first.first, bk = kv.first.second;
        const int bm = gemm_tile_bm(tile), bn = gemm_tile_bn(tile);
        std::vector<GemmProb> probs;
        std::vector<int> off{0};
        for (int q : kv.second) {
            const GemmDesc& g = list[q];
            TN_REQUIRE(gemm_tile_ok(tile, g.tin, g.M, g.N, g.K), TN_ERR_ARG, "gemm_group: tile does not fit");
            probs.push_back(GemmProb{g.A, g.B, g.C, g.lda, g.ldb, g.ldc, g.M, g.N, g.K, g.alpha, g.accumulate});
            off.push_back(off.back() + (g.M / bm) * (g.N / bn));
        }
        const GemmProb* dp = stage(b, probs);
        const int* doff = stage(b, off);
        out.push_back(GroupLaunch{list[0].tin, list[0].tout, list[0].a_kmajor, list[0].b_kmajor, tile, bk, dp, doff,
                                  (int)probs.size(), off.back()});
    }
    return out;
}

void gemm_group_launch(tn_batch* b, const std::vector<GroupLaunch>& L)
{
    for (const GroupLaunch& g : L)
        TN_HIP(gemm_grouped(g.tin, g.tout, g.akm, g.bkm, g.tile, g.bk, g.probs, g.off, g.nprob, g.total,
                            b->ctx->stream));
}

void gemm_group(tn_batch* b, const std::vector<GemmDesc>& list)
{
    std::vector<GroupLaunch> L = gemm_group_stage(b, list);
    flush(b);
    gemm_group_launch(b, L);
}

// one active chain of a round
struct Slot {
    int q;      // chain index in the batch
    int site;   // site being updated
    HeffOp op;  // FP32 operator of that site
    long N;     // padded vector length
};

// y_s = Heff_s x_s for every slot (FP32 contractions, FP64 in/out).  An
// identity environment channel of a slot (op.wl_id / op.wr_id) is applied as
// a copy, as in heff_apply: stage 1 skips the right one, the mixing stage
// reads x for it and writes the left one's slice into y, stage 3 skips the
// left one and accumulates.
struct HeffLaunch {
    const ConvJob* conv = nullptr;
    int nconv = 0;
    long nmax = 0;
    std::vector<GroupLaunch> g1, g3;
    const MixJob* mix = nullptr;
    int nmix = 0, Xmax = 0, Bnmax = 0;
};

HeffLaunch heff_batch_stage(tn_batch* b, const std::vector<Slot>& S, const std::vector<const double*>& x,
                            const std::vector<double*>& y)
{
    std::vector<ConvJob> cj;
    std::vector<GemmDesc> g1s, g3s;
    std::vector<MixJob> mj;
    long nmax = 0;
    int Xmax = 0, Bnmax = 0;
    for (size_t k = 0; k < S.size(); k++) {
        const HeffOp& op = S[k].op;
        const int q = S[k].q, p = op.p, D1p = op.D1p, D2p = op.D2p, Dw1 = op.Dw1, Dw2 = op.Dw2;
        const bool idl = op.wl_id >= 0, idr = op.wr_id >= 0;
        const int nR = Dw2 - (idr ? 1 : 0), r0 = op.wr_id == 0 ? 1 : 0;
        const int nL = Dw1 - (idl ? 1 : 0), l0 = op.wl_id == 0 ? 1 : 0;
        float* xl = b->xlo[q].as<float>();
        cj.push_back(ConvJob{x[k], xl, op.n()});
        nmax = std::max(nmax, op.n());
        GemmDesc g1{};
        g1.tin = T_F32; g1.tout = T_F32; g1.a_kmajor = true; g1.b_kmajor = true;
        g1.M = D1p * p; g1.N = nR * D2p; g1.K = D2p; g1.batch = 1;
        g1.A = xl; g1.lda = D2p; g1.B = (const float*)op.Rlo + (long)r0 * D2p * D2p; g1.ldb = D2p;
        g1.C = b->T1[q].p; g1.ldc = (long)nR * D2p; g1.alpha = 1.0f;
        g1s.push_back(g1);
        MixJob m{b->T1[q].p, (long)p * nR * D2p, b->T2[q].p, (long)p * D2p, D1p, D2p, op.mix->t, nullptr, 0, nullptr, 0};
        if (idl || idr) {
            TN_REQUIRE(op.mix->op == MIX_HEFF && op.mix->t.esel && op.mix->wl == op.wl_id && op.mix->wr == op.wr_id,
                       TN_ERR_STATE,
                       "heff_batch: mixing table does not match the identity channels");
            m.in2 = xl;
            m.isx2 = (long)p * D2p;
            m.out2 = y[k];
            m.osx2 = (long)p * D2p;
        }
        mj.push_back(m);
        Xmax = std::max(Xmax, D1p);
        Bnmax = std::max(Bnmax, D2p);
        GemmDesc g3{};
        g3.tin = T_F32; g3.tout = T_F64; g3.a_kmajor = true; g3.b_kmajor = false;
        g3.M = D1p; g3.N = p * D2p; g3.K = nL * D1p; g3.batch = 1;
        g3.A = (const float*)op.Llo + (long)l0 * D1p; g3.lda = (long)Dw1 * D1p; g3.B = b->T2[q].p;
        g3.ldb = (long)p * D2p;
        g3.C = y[k]; g3.ldc = (long)p * D2p; g3.alpha = 1.0f; g3.accumulate = idl ? 1 : 0;
        g3s.push_back(g3);
    }
    HeffLaunch h;
    h.conv = stage(b, cj);
    h.nconv = (int)cj.size();
    h.nmax = nmax;
    h.g1 = gemm_group_stage(b, g1s);
    h.mix = stage(b, mj);
    h.nmix = (int)mj.size();
    h.Xmax = Xmax;
    h.Bnmax = Bnmax;
    h.g3 = gemm_group_stage(b, g3s);  // the accumulate flag travels per problem
    return h;
}

// conv_done: the FP32 copies of x were written by the caller (fused with the
// Lanczos scaling)
void heff_batch_launch(tn_batch* b, const HeffLaunch& h, bool conv_done = false)
{
    hipStream_t s = b->ctx->stream;
    if (!conv_done) TN_HIP(bconvert(h.conv, h.nconv, h.nmax, s));
    gemm_group_launch(b, h.g1);
    TN_HIP(bwmix(T_F32, h.mix, h.nmix, h.Xmax, h.Bnmax, s));
    gemm_group_launch(b, h.g3);
}

// squared norms of the vectors v_s (length n_s), read back
std::vector<double> norms2(tn_batch* b, const std::vector<const double*>& v, const std::vector<long>& n)
{
    std::vector<DotJob> dj;
    long nmax = 1;
    double* out = b->coef.as<double>();
    for (size_t k = 0; k < v.size(); k++) {
        dj.push_back(DotJob{v[k], n[k], 1, v[k], n[k], out + k});
        nmax = std::max(nmax, n[k]);
    }
    TN_HIP(bmdot(push(b, dj), (int)dj.size(), 1, nmax, b->part.as<double>(), b->ctx->stream));
    std::vector<double> r(v.size());
    readback(b, r.data(), out, v.size());
    return r;
}

// Lockstep Lanczos (the algorithm of tn::lanczos: three-term step, one
// Gram-Schmidt pass against all Krylov vectors, host tridiagonal solve) on
// every slot; the Ritz vector replaces the slot's site tensor.
//
// partial: Simon's partial re-orthogonalisation as in tn::lanczos (eps_op the
// operator's unit roundoff).  Each step works against v_{j-1}, v_j only; after
// the read-back of iteration j the host propagates each chain's omega
// estimates and, before launching iteration j+1, queues a pass against the
// whole basis for the chains above sqrt(eps_op).  Coefficient slots per chain
// (h = coef + k*stride): h[0] alpha dot, h[1..2] local dots, h[3 + (j & 1)]
// |w_j|^2 (the pass overwrites it with the corrected value, which the device
// reads as beta_j in iteration j+1), h[8..8+j] the pass's dots.  A chain with
// a pass at j gets its corrected alpha_j, beta_j with the read-back of j+1,
// so its tridiagonal solve of j is skipped.
void lanczos_batch(tn_batch* b, const std::vector<Slot>& S, double tol, int max_restarts, std::vector<double>& e0,
                   std::vector<double>& e1, tn_sweep_stats* st, bool partial, double eps_op)
{
    tn_ctx* c = b->ctx;
    hipStream_t s = c->stream;
    const int ns = (int)S.size(), kmax = b->kmax;
    const size_t str = b->stride;
    TN_REQUIRE(!partial || str >= (size_t)kmax + 8, TN_ERR_STATE, "batch Lanczos: coefficient stride too small");
    double* coef = b->coef.as<double>();
    e0.assign(ns, 0.0);
    e1.assign(ns, 0.0);
    std::vector<int> todo(ns);
    for (int k = 0; k < ns; k++) todo[k] = k;
    std::vector<std::vector<double>> al(ns, std::vector<double>(kmax)), be(ns, std::vector<double>(kmax)),
        z(ns, std::vector<double>(kmax));
    std::vector<int> kk(ns, 0);
    // partial re-orthogonalisation state per chain
    const double eta = std::sqrt(eps_op);
    std::vector<std::vector<double>> om_old, om, om_new;
    if (partial) {
        om_old.assign(ns, std::vector<double>(kmax + 2, 0.0));
        om.assign(ns, std::vector<double>(kmax + 2, 0.0));
        om_new.assign(ns, std::vector<double>(kmax + 2, 0.0));
    }
    std::vector<char> force(ns, 0), pend(ns, 0), full(ns, 0);
    std::vector<double> tnorm(ns, 0.0);
    auto V = [&](int k) { return b->kry[S[k].q].as<double>(); };
    auto site = [&](int k) { return b->ch[S[k].q]->A[S[k].site]; };
    for (int restart = 0; restart <= max_restarts && !todo.empty(); restart++) {
        // v0 = x / |x|
        {
            std::vector<const double*> xv;
            std::vector<long> nv;
            for (int k : todo) {
                xv.push_back(site(k));
                nv.push_back(S[k].N);
            }
            std::vector<double> n2 = norms2(b, xv, nv);
            std::vector<ScaleJob> sj;
            long nmax = 1;
            for (size_t t = 0; t < todo.size(); t++) {
                const int k = todo[t];
                TN_REQUIRE(n2[t] > 0.0 && std::isfinite(n2[t]), TN_ERR_STATE, "batch Lanczos: zero start vector");
                sj.push_back(ScaleJob{1.0 / std::sqrt(n2[t]), site(k), V(k), S[k].N});
                nmax = std::max(nmax, S[k].N);
            }
            TN_HIP(bscale(push(b, sj), (int)sj.size(), nmax, s));
        }
        for (int k : todo) {
            force[k] = pend[k] = full[k] = 0;
            tnorm[k] = 0.0;
            if (partial) om[k][0] = 1.0;
        }
        // Pipelined lockstep iterations.  After the read-back of iteration j
        // the host launches iteration j+1 for every chain still active
        // (v_{j+1} = w / beta_j with beta_j taken from the device copy of
        // |w|^2), then solves the tridiagonal problems of iteration j while the
        // device works.  A chain found converged at j has run one iteration
        // more than it needs; its results are ignored.
        std::vector<int> act = todo;
        std::vector<double> resid(ns, 0.0);
        std::vector<char> fin(ns, 0);
        // slot of |w_j|^2 in a chain's coefficients
        auto nslot = [&](int j) { return partial ? 3 + (j & 1) : j + 2; };
        auto launch_iter = [&](int j, const std::vector<int>& A) {
            std::vector<Slot> sa;
            std::vector<const double*> xin;
            std::vector<double*> yout;
            std::vector<ScaleConvJob> sc;
            long nmax = 1;
            for (int k : A) {
                const long N = S[k].N;
                double* vj = V(k) + (long)j * N;
                double* h = coef + (size_t)k * str;
                sa.push_back(S[k]);
                xin.push_back(vj);
                yout.push_back(V(k) + (long)(j + 1) * N);
                nmax = std::max(nmax, N);
                // v_j = w_{j-1} / |w_{j-1}| (|w_{j-1}|^2 of iteration j-1) and
                // its FP32 copy for the Heff stages, in one pass
                if (j > 0) sc.push_back(ScaleConvJob{0.0, vj, vj, b->xlo[S[k].q].as<float>(), N, h + nslot(j - 1)});
            }
            // every descriptor of the iteration in one transfer: the scaling,
            // the Heff stages, then the Lanczos step (alpha = v_j . w; three-term
            // step fused with the Gram-Schmidt dots; Gram-Schmidt update fused
            // with the norm, the kernels of tn::lanczos)
            const ScaleConvJob* dsc = sc.empty() ? nullptr : stage(b, sc);
            HeffLaunch hl = heff_batch_stage(b, sa, xin, yout);
            std::vector<DotJob> d1;
            std::vector<StepJob> sj3;
            std::vector<UpdJob> uj;
            // the step works against v_{j-1}, v_j (partial) or every Krylov vector
            const int kl = partial ? std::min(j + 1, 2) : j + 1;
            for (int k : A) {
                double* Vk = V(k);
                const long N = S[k].N;
                double* w = Vk + (long)(j + 1) * N;
                double* h = coef + (size_t)k * str;
                const double* Vl = Vk + (long)(j + 1 - kl) * N;
                d1.push_back(DotJob{Vk + (long)j * N, N, 1, w, N, h});
                sj3.push_back(StepJob{w, Vl, N, kl, h, 0.0, N, h + 1, j > 0 ? h + nslot(j - 1) : nullptr});
                uj.push_back(UpdJob{w, Vl, N, kl, h + 1, N, h + nslot(j)});
            }
            const DotJob* dd1 = stage(b, d1);
            const StepJob* dsj = stage(b, sj3);
            const UpdJob* duj = stage(b, uj);
            flush(b);
            const int na = (int)A.size();
            double* part = b->part.as<double>();
            if (dsc) TN_HIP(bscale_conv(dsc, (int)sc.size(), nmax, s));
            if (st) TN_HIP(hipEventRecord(c->ev[0], s));
            heff_batch_launch(b, hl, dsc != nullptr);
            if (st) TN_HIP(hipEventRecord(c->ev[1], s));
            TN_HIP(bmdot(dd1, na, 1, nmax, part, s));
            TN_HIP(bstep_dots(dsj, na, kl, nmax, part, s));
            TN_HIP(bupdate_norm(duj, na, nmax, part, s));
        };
        // partial: the pass of w_j against v_0..v_j for the chains in F
        auto launch_full = [&](int j, const std::vector<int>& F) {
            std::vector<DotJob> dj;
            std::vector<UpdJob> uj;
            long nmax = 1;
            for (int k : F) {
                double* Vk = V(k);
                const long N = S[k].N;
                double* w = Vk + (long)(j + 1) * N;
                double* h = coef + (size_t)k * str;
                dj.push_back(DotJob{Vk, N, j + 1, w, N, h + 8});
                uj.push_back(UpdJob{w, Vk, N, j + 1, h + 8, N, h + nslot(j)});
                nmax = std::max(nmax, N);
            }
            const DotJob* ddj = stage(b, dj);
            const UpdJob* duj = stage(b, uj);
            flush(b);
            double* part = b->part.as<double>();
            TN_HIP(bmdot(ddj, (int)F.size(), j + 1, nmax, part, s));
            TN_HIP(bupdate_norm(duj, (int)F.size(), nmax, part, s));
        };
        double t0 = now_ms();
        if (kmax > 0 && !act.empty()) launch_iter(0, act);
        for (int j = 0; j < kmax && !act.empty(); j++) {
            readback(b, b->hcoef.data(), coef, (size_t)ns * str);
            if (st) {
                float ms = 0;
                TN_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                st->heff_ms += ms;
                st->lanczos_ms += now_ms() - t0;
            }
            t0 = now_ms();
            const int kl = partial ? std::min(j + 1, 2) : j + 1;
            if (partial) {
                // chains whose w_{j-1} had a pass against the whole basis: its dot
                // with v_{j-1} and the corrected |w_{j-1}|^2.  When the pass left
                // only rounding (an invariant subspace), v_j was noise: the chain
                // ends at iteration j-1 and the results of j are dropped.
                std::vector<int> keep;
                for (int k : act) {
                    if (pend[k]) {
                        const double* h = b->hcoef.data() + (size_t)k * str;
                        al[k][j - 1] += h[8 + (j - 1)];
                        const double bc = std::sqrt(std::max(h[nslot(j - 1)], 0.0));
                        pend[k] = 0;
                        if (bc < 1e-10 * be[k][j - 1] || bc < 1e-14 * std::max(1.0, std::fabs(e0[k]))) {
                            be[k][j - 1] = 0.0;
                            tridiag_lowest(j, al[k].data(), be[k].data(), &e0[k], &e1[k], z[k].data(), true);
                            kk[k] = j;
                            resid[k] = 0.0;
                            continue;
                        }
                        be[k][j - 1] = bc;
                    }
                    keep.push_back(k);
                }
                act.swap(keep);
                if (act.empty()) break;
            }
            for (int k : act) {
                const double* h = b->hcoef.data() + (size_t)k * str;
                al[k][j] = h[0] + h[kl];
                be[k][j] = std::sqrt(std::max(h[nslot(j)], 0.0));
            }
            std::vector<int> F;
            if (partial) {
                // omega_{j+1,.} of every active chain; a pass for those above eta
                // (none in the last iteration: w_j is not used)
                for (int k : act) {
                    std::vector<double>&on = om_new[k], &oc = om[k], &oo = om_old[k];
                    const double* a = al[k].data();
                    const double* bb = be[k].data();
                    tnorm[k] = std::max(tnorm[k], std::fabs(a[j]) + bb[j] + (j > 0 ? bb[j - 1] : 0.0));
                    const double noise = eps_op * tnorm[k];
                    double wmax = 0.0;
                    if (bb[j] > 0.0) {
                        for (int q = 0; q + 1 < j; q++) {
                            double t = bb[q] * oc[q + 1] + (a[q] - a[j]) * oc[q] - bb[j - 1] * oo[q];
                            if (q > 0) t += bb[q - 1] * oc[q - 1];
                            t = (t + std::copysign(noise, t)) / bb[j];
                            on[q] = t;
                            wmax = std::max(wmax, std::fabs(t));
                        }
                        for (int q = std::max(j - 1, 0); q <= j; q++) on[q] = noise / bb[j];
                    }
                    on[j + 1] = 1.0;
                    full[k] = 0;
                    if (j > 0 && j + 1 < kmax && (force[k] || wmax > eta || !std::isfinite(wmax))) {
                        for (int q = 0; q <= j; q++) on[q] = bb[j] > 0.0 ? noise / bb[j] : 0.0;
                        force[k] = !force[k];
                        full[k] = 1;
                        F.push_back(k);
                    } else {
                        force[k] = 0;
                    }
                    std::swap(oo, oc);
                    std::swap(oc, on);
                }
                if (!F.empty()) {
                    launch_full(j, F);
                    if (st) st->n_full_reorth += (int64_t)F.size();
                }
            } else if (st) {
                st->n_full_reorth += (int64_t)act.size();
            }
            if (j + 1 < kmax) launch_iter(j + 1, act);
            // the tridiagonal Ritz problems of the active chains, in parallel
            // with the device and on the host threads.  A chain whose w_j had a
            // pass against the whole basis is tested with the values before the
            // pass (its beta_j is then an upper bound); if it goes on, the
            // corrected alpha_j, beta_j arrive with the next read-back.
            b->pool->run((int)act.size(), [&](int t) {
                const int k = act[t];
                tridiag_lowest(j + 1, al[k].data(), be[k].data(), &e0[k], nullptr, z[k].data(), false);
                kk[k] = j + 1;
                resid[k] = std::fabs(be[k][j] * z[k][j]);
                const bool done = resid[k] < tol || be[k][j] < 1e-14 * std::max(1.0, std::fabs(e0[k]));
                fin[k] = done ? 1 : 0;
                pend[k] = full[k] && !done;
                if (done || j + 1 == kmax)  // the second Ritz value is only reported
                    tridiag_lowest(j + 1, al[k].data(), be[k].data(), &e0[k], &e1[k], z[k].data(), true);
            });
            std::vector<int> next;
            for (int k : act) {
                if (st) {
                    st->n_matvec++;
                    const double D1 = b->ch[S[k].q]->D1[S[k].site], D2 = b->ch[S[k].q]->D2[S[k].site];
                    st->heff_flops += S[k].op.flops_alg((int)D1, (int)D2);
                }
                TN_REQUIRE(std::isfinite(e0[k]), TN_ERR_STATE, "batch Lanczos: non-finite Ritz value");
                if (!fin[k]) next.push_back(k);
            }
            act.swap(next);
        }
        // Ritz vectors x = V z, normalised, for every chain of this cycle
        {
            std::vector<double> zh((size_t)ns * str, 0.0);
            for (int k : todo)
                for (int i = 0; i < kk[k]; i++) zh[(size_t)k * str + i] = z[k][i];
            h2d(c, coef, zh.data(), zh.size() * 8);
            ring_reset(b);
            std::vector<GemvJob> gj;
            std::vector<const double*> xv;
            std::vector<long> nv;
            long nmax = 1;
            for (int k : todo) {
                gj.push_back(GemvJob{V(k), S[k].N, kk[k], coef + (size_t)k * str, 1.0, 0.0, site(k), S[k].N});
                xv.push_back(site(k));
                nv.push_back(S[k].N);
                nmax = std::max(nmax, S[k].N);
            }
            TN_HIP(bgemv(push(b, gj), (int)gj.size(), nmax, s));
            std::vector<double> n2 = norms2(b, xv, nv);
            std::vector<ScaleJob> sj;
            for (size_t t = 0; t < todo.size(); t++)
                sj.push_back(ScaleJob{1.0 / std::sqrt(n2[t]), xv[t], (double*)xv[t], nv[t]});
            TN_HIP(bscale(push(b, sj), (int)sj.size(), nmax, s));
        }
        // chains still above the tolerance restart from their Ritz vector
        todo = act;
    }
    if (st) {
        st->n_solve_capped += (int64_t)todo.size();
        st->n_solve_converged += (int64_t)(ns - (int)todo.size());
    }
}

// One CholeskyQR pass X_s = Q_s L_s^T on every slot (X: m x n row-major, ld
// ldx): Gram matrix, blocked Cholesky with batched diagonal blocks,
// recursive-doubling inverse, Q = X L^-T.
struct QrSlot {
    int q, m, n, nl;
    const double* X;
    long ldx;
    double* Q;
    double* L;
};

void qr_pass_batch(tn_batch* b, const std::vector<QrSlot>& P)
{
    hipStream_t s = b->ctx->stream;
    std::vector<GemmDesc> gram, apply;
    std::vector<ScaleJob> zero;
    long zmax = 1;
    for (const QrSlot& t : P) {
        GemmDesc g{};
        g.tin = T_F64; g.tout = T_F64; g.a_kmajor = false; g.b_kmajor = false;
        g.M = t.n; g.N = t.n; g.K = t.m; g.batch = 1;
        g.A = t.X; g.lda = t.ldx; g.B = t.X; g.ldb = t.ldx; g.C = b->qG[t.q].p; g.ldc = t.n; g.alpha = 1.0f;
        gram.push_back(g);
        const long nn = (long)t.n * t.n;
        zero.push_back(ScaleJob{0.0, t.L, t.L, nn});
        zero.push_back(ScaleJob{0.0, b->qLinv[t.q].as<double>(), b->qLinv[t.q].as<double>(), nn});
        zmax = std::max(zmax, nn);
    }
    gemm_group(b, gram);
    TN_HIP(bscale(push(b, zero), (int)zero.size(), zmax, s));
    // identity on the zero padding of every Gram matrix
    {
        std::vector<DiagJob> dj;
        int lenmax = 0;
        for (const QrSlot& t : P)
            if (t.nl < t.n) {
                dj.push_back(DiagJob{b->qG[t.q].as<double>(), t.n, t.nl, t.n, 1.0});
                lenmax = std::max(lenmax, t.n - t.nl);
            }
        if (!dj.empty()) TN_HIP(badd_diag(push(b, dj), (int)dj.size(), lenmax, s));
    }
    int maxnb = 0;
    for (const QrSlot& t : P) maxnb = std::max(maxnb, (t.n + 63) / 64);
    for (int kb = 0; kb < maxnb; kb++) {
        std::vector<CholJob> cj;
        std::vector<GemmDesc> panel, trail;
        for (const QrSlot& t : P) {
            const int k0 = kb * 64;
            if (k0 >= t.n) continue;
            const int nbk = std::min(64, t.n - k0), r0 = k0 + nbk, rem = t.n - r0;
            double* G = b->qG[t.q].as<double>();
            double* Li = b->qLinv[t.q].as<double>();
            cj.push_back(CholJob{G, t.n, k0, nbk, t.L, t.n, Li, t.n, b->flags.as<int>() + t.q});
            if (rem <= 0) continue;
            GemmDesc p{};
            p.tin = T_F64; p.tout = T_F64; p.a_kmajor = true; p.b_kmajor = true;
            p.M = rem; p.N = nbk; p.K = nbk; p.batch = 1;
            p.A = G + (size_t)r0 * t.n + k0; p.lda = t.n;
            p.B = Li + (size_t)k0 * t.n + k0; p.ldb = t.n;
            p.C = t.L + (size_t)r0 * t.n + k0; p.ldc = t.n; p.alpha = 1.0f;
            panel.push_back(p);
            GemmDesc u{};
            u.tin = T_F64; u.tout = T_F64; u.a_kmajor = true; u.b_kmajor = true;
            u.M = rem; u.N = rem; u.K = nbk; u.batch = 1;
            u.A = t.L + (size_t)r0 * t.n + k0; u.lda = t.n;
            u.B = t.L + (size_t)r0 * t.n + k0; u.ldb = t.n;
            u.C = G + (size_t)r0 * t.n + r0; u.ldc = t.n; u.alpha = -1.0f; u.accumulate = 1;
            trail.push_back(u);
        }
        const CholJob* dcj = stage(b, cj);
        std::vector<GroupLaunch> lp = gemm_group_stage(b, panel), lt = gemm_group_stage(b, trail);
        flush(b);
        TN_HIP(chol_diag_batch(dcj, (int)cj.size(), s));
        gemm_group_launch(b, lp);
        gemm_group_launch(b, lt);
    }
    // L^-1 by recursive doubling of block pairs (as in cholqr_device)
    int maxn = 0;
    for (const QrSlot& t : P) maxn = std::max(maxn, t.n);
    for (int bsz = 64; bsz < maxn; bsz *= 2) {
        std::vector<GemmDesc> xs, ys;
        for (const QrSlot& t : P) {
            if (bsz >= t.n) continue;
            const int Pn = t.n / (2 * bsz), tail = t.n - Pn * 2 * bsz;
            double* Li = b->qLinv[t.q].as<double>();
            double* X = b->qT[t.q].as<double>();
            int slot = 0;
            auto add = [&](size_t o, int s2) {
                double* Xk = X + (size_t)slot * bsz * bsz;
                slot++;
                GemmDesc x{};
                x.tin = T_F64; x.tout = T_F64; x.a_kmajor = true; x.b_kmajor = false;
                x.M = s2; x.N = bsz; x.K = bsz; x.batch = 1;
                x.A = t.L + (o + bsz) * t.n + o; x.lda = t.n;
                x.B = Li + o * t.n + o; x.ldb = t.n;
                x.C = Xk; x.ldc = bsz; x.alpha = 1.0f;
                xs.push_back(x);
                GemmDesc y{};
                y.tin = T_F64; y.tout = T_F64; y.a_kmajor = true; y.b_kmajor = false;
                y.M = s2; y.N = bsz; y.K = s2; y.batch = 1;
                y.A = Li + (o + bsz) * t.n + o + bsz; y.lda = t.n;
                y.B = Xk; y.ldb = bsz;
                y.C = Li + (o + bsz) * t.n + o; y.ldc = t.n; y.alpha = -1.0f;
                ys.push_back(y);
            };
            for (int k = 0; k < Pn; k++) add((size_t)k * 2 * bsz, bsz);
            if (tail > bsz) add((size_t)Pn * 2 * bsz, tail - bsz);
        }
        gemm_group(b, xs);
        gemm_group(b, ys);
    }
    for (const QrSlot& t : P) {
        GemmDesc a{};
        a.tin = T_F64; a.tout = T_F64; a.a_kmajor = true; a.b_kmajor = true;
        a.M = t.m; a.N = t.n; a.K = t.n; a.batch = 1;
        a.A = t.X; a.lda = t.ldx; a.B = b->qLinv[t.q].p; a.ldb = t.n; a.C = t.Q; a.ldc = t.n; a.alpha = 1.0f;
        apply.push_back(a);
    }
    gemm_group(b, apply);
}

// Gram deviation max|Q^T Q - I| (logical block) of every slot, read back
std::vector<double> qr_check_batch(tn_batch* b, const std::vector<QrSlot>& P)
{
    hipStream_t s = b->ctx->stream;
    std::vector<GemmDesc> gram;
    for (const QrSlot& t : P) {
        GemmDesc g{};
        g.tin = T_F64; g.tout = T_F64; g.a_kmajor = false; g.b_kmajor = false;
        g.M = t.n; g.N = t.n; g.K = t.m; g.batch = 1;
        g.A = t.Q; g.lda = t.n; g.B = t.Q; g.ldb = t.n; g.C = b->qG[t.q].p; g.ldc = t.n; g.alpha = 1.0f;
        gram.push_back(g);
    }
    gemm_group(b, gram);
    double* out = b->coef.as<double>();
    std::vector<EyeJob> ej;
    for (size_t k = 0; k < P.size(); k++) ej.push_back(EyeJob{b->qG[P[k].q].as<double>(), P[k].n, P[k].nl, out + k});
    TN_HIP(beye_deviation(push(b, ej), (int)ej.size(), b->part.as<double>(), s));
    std::vector<double> dev(P.size());
    readback(b, dev.data(), out, P.size());
    return dev;
}

// Gauge moves of a round (tn::svd_normalize with gauge = 1, batched): the
// site becomes left- (dir 0) or right-normalised (dir 1) by CholeskyQR2 and
// the neighbour absorbs the triangular factor.  Edge sites and chains whose
// factorisation misses the bound take the single-chain path.
void gauge_batch(tn_batch* b, const std::vector<Slot>& S, int dir, tn_sweep_stats* st)
{
    tn_ctx* c = b->ctx;
    hipStream_t s = c->stream;
    const double tol = 1e-11;
    double t0 = now_ms();
    std::vector<int> bat;
    for (int k = 0; k < (int)S.size(); k++) {
        tn_chain* ch = b->ch[S[k].q];
        const int i = S[k].site, d = ch->d, D1 = ch->D1[i], D2 = ch->D2[i];
        const bool edge = (dir == 0 && i == ch->L - 1) || (dir == 1 && i == 0);
        const bool wide = dir == 0 ? D1 * d < D2 : d * D2 < D1;
        if (edge || wide) svd_normalize(ch, i, dir, st, 1);
        else bat.push_back(k);
    }
    if (bat.empty()) return;
    std::vector<QrSlot> P1, P2;
    std::vector<Copy2Job> xcp;
    int xro = 0, xco = 0;
    for (int k : bat) {
        tn_chain* ch = b->ch[S[k].q];
        const int q = S[k].q, i = S[k].site, d = ch->d;
        const int D1p = pad32(ch->D1[i]), D2p = pad32(ch->D2[i]);
        QrSlot t{};
        t.q = q;
        if (dir == 0) {
            t.m = D1p * d; t.n = D2p; t.nl = ch->D2[i]; t.X = ch->A[i]; t.ldx = D2p;
        } else {
            t.m = d * D2p; t.n = D1p; t.nl = ch->D1[i];
            xcp.push_back(Copy2Job{ch->A[i], (long)d * D2p, 1, D1p, d * D2p, b->qX[q].as<double>(), 1, D1p, D1p,
                                   d * D2p});
            xro = std::max(xro, D1p);
            xco = std::max(xco, d * D2p);
            t.X = b->qX[q].as<double>(); t.ldx = D1p;
        }
        t.Q = b->qQ1[q].as<double>();
        t.L = b->qL[q].as<double>();
        P1.push_back(t);
        QrSlot t2 = t;
        t2.X = t.Q; t2.ldx = t.n; t2.Q = b->qQ[q].as<double>(); t2.L = b->qLk[q].as<double>();
        P2.push_back(t2);
    }
    if (!xcp.empty()) TN_HIP(bcopy2(push(b, xcp), (int)xcp.size(), xro, xco, s));
    {
        std::vector<int> z(b->ch.size(), 0);
        h2d(c, b->flags.p, z.data(), z.size() * sizeof(int));
        ring_reset(b);
    }
    qr_pass_batch(b, P1);
    qr_pass_batch(b, P2);
    // Lacc = L1 L2  (A = Q1 L1^T = Q L2^T L1^T)
    {
        std::vector<GemmDesc> acc;
        for (const QrSlot& t : P1) {
            GemmDesc g{};
            g.tin = T_F64; g.tout = T_F64; g.a_kmajor = true; g.b_kmajor = false;
            g.M = t.n; g.N = t.n; g.K = t.n; g.batch = 1;
            g.A = t.L; g.lda = t.n; g.B = b->qLk[t.q].p; g.ldb = t.n; g.C = b->qLacc[t.q].p; g.ldc = t.n;
            g.alpha = 1.0f;
            acc.push_back(g);
        }
        gemm_group(b, acc);
    }
    std::vector<double> dev = qr_check_batch(b, P2);
    std::vector<int> flag(b->ch.size(), 0);
    d2h(c, flag.data(), b->flags.p, flag.size() * sizeof(int));
    ring_reset(b);
    // accepted slots: site <- Q, neighbour <- R M (left) or M L (right)
    std::vector<GemmDesc> nb0, nb1;
    std::vector<ScaleJob> cp;
    std::vector<Copy2Job> qcp;
    int qro = 0, qco = 0;
    std::vector<int> ok, resc;
    long cmax = 1;
    for (size_t t = 0; t < bat.size(); t++) {
        const int k = bat[t], q = S[k].q, i = S[k].site;
        tn_chain* ch = b->ch[q];
        const int d = ch->d;
        if (flag[q] || !(dev[t] <= tol)) {
            svd_normalize(ch, i, dir, st, 1);
            if (st) st->n_qr_fallback++;
            continue;
        }
        ok.push_back(k);
        const QrSlot& P = P2[t];
        if (dir == 0) {
            const long nA = (long)P.m * P.n;
            cp.push_back(ScaleJob{1.0, P.Q, ch->A[i], nA});
            cmax = std::max(cmax, nA);
            const int j = i + 1, D3p = pad32(ch->D2[j]);
            GemmDesc g{};
            g.tin = T_F64; g.tout = T_F64; g.a_kmajor = false; g.b_kmajor = false;
            g.M = P.n; g.N = d * D3p; g.K = P.n; g.batch = 1;
            g.A = b->qLacc[q].p; g.lda = P.n; g.B = ch->A[j]; g.ldb = (long)d * D3p;
            g.C = b->qQ1[q].p; g.ldc = (long)d * D3p; g.alpha = 1.0f;
            nb0.push_back(g);
            resc.push_back(j);
        } else {
            const int D1p = P.n, D2p = pad32(ch->D2[i]);
            qcp.push_back(Copy2Job{P.Q, 1, P.n, D1p, d * D2p, ch->A[i], (long)d * D2p, 1, D1p, d * D2p});
            qro = std::max(qro, D1p);
            qco = std::max(qco, d * D2p);
            const int j = i - 1, D0p = pad32(ch->D1[j]);
            GemmDesc g{};
            g.tin = T_F64; g.tout = T_F64; g.a_kmajor = true; g.b_kmajor = false;
            g.M = D0p * d; g.N = P.n; g.K = P.n; g.batch = 1;
            g.A = ch->A[j]; g.lda = P.n; g.B = b->qLacc[q].p; g.ldb = P.n;
            g.C = b->qQ1[q].p; g.ldc = P.n; g.alpha = 1.0f;
            nb1.push_back(g);
            resc.push_back(j);
        }
    }
    if (!cp.empty()) TN_HIP(bscale(push(b, cp), (int)cp.size(), cmax, s));
    if (!qcp.empty()) TN_HIP(bcopy2(push(b, qcp), (int)qcp.size(), qro, qco, s));
    if (ok.empty()) return;
    gemm_group(b, nb0);
    gemm_group(b, nb1);
    // neighbour <- product / |product| (rescale_site), all chains at once
    std::vector<const double*> xv;
    std::vector<long> nv;
    for (size_t t = 0; t < ok.size(); t++) {
        tn_chain* ch = b->ch[S[ok[t]].q];
        const int j = resc[t];
        xv.push_back(b->qQ1[S[ok[t]].q].as<double>());
        nv.push_back((long)pad32(ch->D1[j]) * ch->d * pad32(ch->D2[j]));
    }
    std::vector<double> n2 = norms2(b, xv, nv);
    std::vector<ScaleJob> sj;
    long nmax = 1;
    for (size_t t = 0; t < ok.size(); t++) {
        tn_chain* ch = b->ch[S[ok[t]].q];
        const int i = S[ok[t]].site, j = resc[t];
        sj.push_back(ScaleJob{n2[t] > 0 ? 1.0 / std::sqrt(n2[t]) : 1.0, xv[t], ch->A[j], nv[t]});
        nmax = std::max(nmax, nv[t]);
        // Bond::ClearSingularValues
        const int bond = dir == 0 ? i : i - 1;
        ch->lam[bond].assign(dir == 0 ? ch->D2[i] : ch->D1[i], 0.0);
        ch->maxDelta[bond] = 0.0;
        if (st) st->n_qr++;
    }
    TN_HIP(bscale(push(b, sj), (int)sj.size(), nmax, s));
    if (st) {
        batch_sync(b);
        st->svd_ms += now_ms() - t0;
    }
}

// Environment updates of a round (tn::env_update with FP32 contractions),
// grouped; the FP32 copies of the new environments are refreshed too.
void env_batch(tn_batch* b, const std::vector<Slot>& S, int dir, tn_sweep_stats* st)
{
    hipStream_t s = b->ctx->stream;
    double t0 = now_ms();
    std::vector<ConvJob> cj, lo;
    std::vector<GemmDesc> e1s, e3s;
    std::vector<MixJob> mj;
    long nmax = 1, lomax = 1;
    int Xmax = 0, Bnmax = 0;
    for (const Slot& sl : S) {
        tn_chain* ch = b->ch[sl.q];
        const int i = sl.site, d = ch->d, D1p = pad32(ch->D1[i]), D2p = pad32(ch->D2[i]);
        const int Dw1 = ch->dw1[i], Dw2 = ch->dw2[i];
        const long nA = (long)D1p * d * D2p;
        float* xl = b->xlo[sl.q].as<float>();
        cj.push_back(ConvJob{ch->A[i], xl, nA});
        nmax = std::max(nmax, nA);
        if (dir == 0) {
            mix_build(ch->mixL[i], MIX_ENVL, ch->W[i].data(), Dw1, Dw2, d, D1p, D2p, s);
            GemmDesc e1{};
            e1.tin = T_F32; e1.tout = T_F32; e1.a_kmajor = true; e1.b_kmajor = false;
            e1.M = D1p * Dw1; e1.N = d * D2p; e1.K = D1p; e1.batch = 1;
            e1.A = lo_left(ch, TN_FP32, i); e1.lda = D1p; e1.B = xl; e1.ldb = (long)d * D2p;
            e1.C = b->T1[sl.q].p; e1.ldc = (long)d * D2p; e1.alpha = 1.0f;
            e1s.push_back(e1);
            mj.push_back(MixJob{b->T1[sl.q].p, (long)Dw1 * d * D2p, b->T2[sl.q].p, (long)d * Dw2 * D2p, D1p, D2p,
                                ch->mixL[i].t});
            GemmDesc e3{};
            e3.tin = T_F32; e3.tout = T_F64; e3.a_kmajor = false; e3.b_kmajor = false;
            e3.M = D2p; e3.N = Dw2 * D2p; e3.K = D1p * d; e3.batch = 1;
            e3.A = xl; e3.lda = D2p; e3.B = b->T2[sl.q].p; e3.ldb = (long)Dw2 * D2p;
            e3.C = ch->Lenv[i]; e3.ldc = (long)Dw2 * D2p; e3.alpha = 1.0f;
            e3s.push_back(e3);
            const long ne = (long)D2p * Dw2 * D2p;
            lo.push_back(ConvJob{ch->Lenv[i], (float*)ch->lo[TN_FP32].L[i], ne});
            lomax = std::max(lomax, ne);
        } else {
            mix_build(ch->mixR[i], MIX_ENVR, ch->W[i].data(), Dw1, Dw2, d, D1p, D2p, s);
            GemmDesc f1{};
            f1.tin = T_F32; f1.tout = T_F32; f1.a_kmajor = true; f1.b_kmajor = true;
            f1.M = D1p * d; f1.N = Dw2 * D2p; f1.K = D2p; f1.batch = 1;
            f1.A = xl; f1.lda = D2p; f1.B = lo_right(ch, TN_FP32, i); f1.ldb = D2p;
            f1.C = b->T1[sl.q].p; f1.ldc = (long)Dw2 * D2p; f1.alpha = 1.0f;
            e1s.push_back(f1);
            mj.push_back(MixJob{b->T1[sl.q].p, (long)d * Dw2 * D2p, b->T2[sl.q].p, (long)Dw1 * d * D2p, D1p, D2p,
                                ch->mixR[i].t});
            for (int w = 0; w < Dw1; w++) {
                GemmDesc f3{};
                f3.tin = T_F32; f3.tout = T_F64; f3.a_kmajor = true; f3.b_kmajor = true;
                f3.M = D1p; f3.N = D1p; f3.K = d * D2p; f3.batch = 1;
                f3.A = xl; f3.lda = (long)d * D2p;
                f3.B = b->T2[sl.q].as<float>() + (long)w * d * D2p; f3.ldb = (long)Dw1 * d * D2p;
                f3.C = ch->Renv[i] + (long)w * D1p * D1p; f3.ldc = D1p; f3.alpha = 1.0f;
                e3s.push_back(f3);
            }
            const long ne = (long)Dw1 * D1p * D1p;
            lo.push_back(ConvJob{ch->Renv[i], (float*)ch->lo[TN_FP32].R[i], ne});
            lomax = std::max(lomax, ne);
        }
        Xmax = std::max(Xmax, D1p);
        Bnmax = std::max(Bnmax, D2p);
        if (st) {
            st->n_env++;
            const double D1 = ch->D1[i], D2 = ch->D2[i];
            st->env_flops +=
                2.0 * D1 * Dw1 * D1 * d * D2 + 2.0 * Dw1 * Dw2 * d * d * D1 * D2 + 2.0 * D2 * Dw2 * D2 * D1 * d;
        }
    }
    if (cj.empty()) return;
    TN_HIP(bconvert(push(b, cj), (int)cj.size(), nmax, s));
    gemm_group(b, e1s);
    TN_HIP(bwmix(T_F32, push(b, mj), (int)mj.size(), Xmax, Bnmax, s));
    std::vector<GemmDesc> e3a, e3b;
    for (auto& g : e3s) (g.a_kmajor ? e3b : e3a).push_back(g);
    gemm_group(b, e3a);
    gemm_group(b, e3b);
    TN_HIP(bconvert(push(b, lo), (int)lo.size(), lomax, s));
    if (st) {
        batch_sync(b);
        st->env_ms += now_ms() - t0;
    }
}

}  // namespace

extern "C" {

int tn_batch_create(tn_ctx* c, int n, tn_chain* const* chains, int max_krylov, tn_batch** out)
{
    if (!c || n <= 0 || !chains || !out) return TN_ERR_ARG;
    tn_batch* b = nullptr;
    try {
        b = new tn_batch();
        b->ctx = c;
        b->kmax = std::max(2, std::min(max_krylov, 200));
        b->stride = (size_t)b->kmax + 8;
        long nmax = 1;
        for (int q = 0; q < n; q++) {
            TN_REQUIRE(chains[q] && chains[q]->ctx == c, TN_ERR_ARG, "tn_batch_create: chain of another context");
            b->ch.push_back(chains[q]);
        }
        const size_t nb = b->ch.size();
        for (auto* v : {&b->kry, &b->T1, &b->T2, &b->xlo, &b->qX, &b->qQ1, &b->qQ, &b->qG, &b->qL, &b->qLk,
                        &b->qLinv, &b->qT, &b->qLacc})
            v->resize(nb);
        for (size_t q = 0; q < nb; q++) {
            tn_chain* ch = b->ch[q];
            const size_t Dp = ch->Dcapp, site = Dp * ch->d * Dp;
            nmax = std::max(nmax, (long)site);
            b->kry[q].ensure((size_t)(b->kmax + 1) * site * 8);
            b->T1[q].ensure(Dp * ch->d * ch->Dwmax * Dp * 4);
            b->T2[q].ensure(Dp * ch->d * ch->Dwmax * Dp * 4);
            b->xlo[q].ensure(site * 4);
            for (auto* v : {&b->qX, &b->qQ1, &b->qQ}) (*v)[q].ensure(site * 8);
            for (auto* v : {&b->qG, &b->qL, &b->qLk, &b->qLinv, &b->qT, &b->qLacc}) (*v)[q].ensure(Dp * Dp * 8);
            ensure_lo(ch, TN_FP32);
        }
        b->coef.ensure(nb * b->stride * 8 + 4096);
        b->hcoef.assign(nb * b->stride, 0.0);
        b->part.ensure((size_t)std::max(bmdot_part_size((int)nb, b->kmax + 1, nmax), (long)nb * 1024) * 8);
        b->flags.ensure(nb * sizeof(int) + 256);
        b->ring_cap = 8u << 20;
        TN_HIP(hipHostMalloc((void**)&b->ring_h, b->ring_cap, hipHostMallocDefault));
        b->ring_d.ensure(b->ring_cap);
        const unsigned hw = std::thread::hardware_concurrency();
        b->pool = new Pool((int)std::min(7u, hw > 1 ? hw - 1 : 0u));
        *out = b;
        return TN_OK;
    } catch (const TnError& e) {
        c->err = e.msg;
        if (b) tn_batch_destroy(b);
        return e.code;
    }
}

int tn_batch_destroy(tn_batch* b)
{
    if (!b) return TN_OK;
    if (b->ctx && b->ctx->stream) (void)hipStreamSynchronize(b->ctx->stream);
    for (auto* v : {&b->kry, &b->T1, &b->T2, &b->xlo, &b->qX, &b->qQ1, &b->qQ, &b->qG, &b->qL, &b->qLk, &b->qLinv,
                    &b->qT, &b->qLacc})
        for (auto& x : *v) x.release();
    b->coef.release();
    b->part.release();
    b->flags.release();
    b->ring_d.release();
    if (b->ring_h) (void)hipHostFree(b->ring_h);
    delete b->pool;
    delete b;
    return TN_OK;
}

int tn_batch_sweep(tn_batch* b, int dir, const tn_sweep_opts* opts, tn_sweep_stats* st, double* energies)
{
    if (!b || (dir != 0 && dir != 1)) return TN_ERR_ARG;
    tn_ctx* c = b->ctx;
    try {
        tn_sweep_opts o;
        tn_sweep_opts_default(&o);
        if (opts) o = *opts;
        TN_REQUIRE(o.precision == TN_FP32 && o.env_precision == TN_FP32, TN_ERR_ARG,
                   "tn_batch_sweep: the batched path contracts in FP32 (precision = env_precision = TN_FP32)");
        TN_REQUIRE(o.max_krylov >= 2 && o.max_krylov <= b->kmax, TN_ERR_ARG,
                   "tn_batch_sweep: max_krylov outside [2, the batch's]");
        const int nch = (int)b->ch.size();
        hipStream_t s = c->stream;
        int Lmax = 0;
        for (auto* ch : b->ch) {
            Lmax = std::max(Lmax, ch->L);
            // only the FP32 environment copies are kept current here
            if (ch->lo[TN_BF16].active) {
                ch->lo[TN_BF16].buf.release();
                ch->lo[TN_BF16].active = false;
            }
            ensure_lo(ch, TN_FP32);
        }
        const double t0 = now_ms();
        const int saved_kmax = b->kmax;
        b->kmax = o.max_krylov;
        std::vector<double> last(nch, 0.0);
        for (int t = 0; t < Lmax; t++) {
            std::vector<Slot> S;
            for (int q = 0; q < nch; q++) {
                tn_chain* ch = b->ch[q];
                if (t >= ch->L) continue;
                const int i = dir == 0 ? t : ch->L - 1 - t;
                Slot sl;
                sl.q = q;
                sl.site = i;
                S.push_back(sl);
            }
            // identity environment channels of every slot (first / last channel
            // of L and R against the identity, one launch, one read back)
            std::vector<int> wl(S.size(), -1), wr(S.size(), -1);
            if (o.skip_identity) {
                std::vector<EyeJob> ej;
                double* dv = b->coef.as<double>();
                for (size_t k = 0; k < S.size(); k++) {
                    tn_chain* ch = b->ch[S[k].q];
                    const int i = S[k].site, D1 = ch->D1[i], D2 = ch->D2[i], D1p = pad32(D1), D2p = pad32(D2);
                    const int Dw1 = ch->dw1[i], Dw2 = ch->dw2[i];
                    const double* L = i > 0 ? ch->Lenv[i - 1] : ch->bL;
                    const double* R = i < ch->L - 1 ? ch->Renv[i + 1] : ch->bR;
                    ej.push_back(EyeJob{L, (long)Dw1 * D1p, D1, dv + 4 * k});
                    ej.push_back(EyeJob{L + (long)(Dw1 - 1) * D1p, (long)Dw1 * D1p, D1, dv + 4 * k + 1});
                    ej.push_back(EyeJob{R, D2p, D2, dv + 4 * k + 2});
                    ej.push_back(EyeJob{R + (long)(Dw2 - 1) * D2p * D2p, D2p, D2, dv + 4 * k + 3});
                }
                TN_HIP(beye_deviation(push(b, ej), (int)ej.size(), b->part.as<double>(), s));
                std::vector<double> h(4 * S.size());
                readback(b, h.data(), dv, h.size());
                auto ok = [](double v) { return std::isfinite(v) && v <= kIdentTolLow; };
                for (size_t k = 0; k < S.size(); k++) {
                    tn_chain* ch = b->ch[S[k].q];
                    const int i = S[k].site, Dw1 = ch->dw1[i], Dw2 = ch->dw2[i];
                    if (Dw1 >= 2) wl[k] = ok(h[4 * k]) ? 0 : (ok(h[4 * k + 1]) ? Dw1 - 1 : -1);
                    if (Dw2 >= 2) wr[k] = ok(h[4 * k + 2]) ? 0 : (ok(h[4 * k + 3]) ? Dw2 - 1 : -1);
                    if (st) {
                        st->n_ident_left += wl[k] >= 0;
                        st->n_ident_right += wr[k] >= 0;
                    }
                }
            }
            for (size_t k = 0; k < S.size(); k++) {
                S[k].op = make_heff(b->ch[S[k].q], S[k].site, TN_FP32, 0, wl[k], wr[k], MIX_HEFF);
                S[k].N = S[k].op.n();
            }
            std::vector<double> e0, e1;
            lanczos_batch(b, S, o.eig_tol, o.max_restarts, e0, e1, st, o.reorth == 1, unit_roundoff(TN_FP32));
            for (size_t k = 0; k < S.size(); k++) {
                tn_chain* ch = b->ch[S[k].q];
                const int i = S[k].site;
                ch->IterDE[i] = e0[k] - ch->Emin[i];
                ch->Emin[i] = e0[k];
                ch->GapE[i] = e1[k] - 1.0;  // MPSSite_Variational.C:29 as written
                last[S[k].q] = e0[k];
            }
            gauge_batch(b, S, dir, st);
            env_batch(b, S, dir, st);
        }
        batch_sync(b);
        b->kmax = saved_kmax;
        if (st) st->wall_ms += now_ms() - t0;
        if (energies)
            for (int q = 0; q < nch; q++) energies[q] = last[q];
        return TN_OK;
    } catch (const TnError& e) {
        c->err = e.msg;
        return e.code;
    }
}

}  // extern "C"
