// tn_chain.hip -- device-resident MPS and the variational sweep behind the C ABI.
//
// Mirrors the reference's sweep bookkeeping (TensorNetworksImp/MPS/MPSImp_Variational.C,
// MPSSite_Normalization.C, MPSSite_Contractions.C, Bond.C, MPSImp.C): thin SVD
// gauge moves with bond dimensions min(rows, cols), edge rescaling, bond
// singular values for the convergence test, zero-padded bond growth.
// Device layouts (bond dims padded to multiples of 32, padding kept at zero):
//   site  A[a][m][b]   [D1p][d][D2p]
//   left  env after site i   [D2p][Dw2][D2p]  (bra, w, ket)
//   right env from site i on [Dw1][D1p][D1p]  (w, bra, ket)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "tn_chain.h"

using namespace tn;


namespace tn {

int fail(tn_ctx* c, const TnError& e)
{
    if (c) {
        c->err = e.msg;
        c->heff_pending = nullptr;  // the caller's stats may not outlive the call
    }
    return e.code;
}

#define ABI_BEGIN try {
#define ABI_END(ctx)                                                  \
    }                                                                 \
    catch (const TnError& e) { return fail(ctx, e); }                 \
    catch (const std::exception& e) { return fail(ctx, TnError{TN_ERR_STATE, e.what()}); }

inline long psi_ref(int d, int D1, int m, int i1, int i2) { return m + (long)d * (i1 + (long)D1 * i2); }

double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// --- host <-> device site transfers (reference layout on the host) -------
void upload_site(tn_chain* ch, int i, int D1, int D2, const double* Mref)
{
    const int d = ch->d, D1p = pad32(D1), D2p = pad32(D2);
    TN_REQUIRE(D1 <= ch->Dcap && D2 <= ch->Dcap, TN_ERR_ARG, "site dims exceed Dcap");
    std::vector<double> buf((size_t)D1p * d * D2p, 0.0);
    for (int a = 0; a < D1; a++)
        for (int m = 0; m < d; m++)
            for (int b = 0; b < D2; b++) buf[((size_t)a * d + m) * D2p + b] = Mref[psi_ref(d, D1, m, a, b)];
    TN_HIP(hipMemcpyAsync(ch->A[i], buf.data(), buf.size() * 8, hipMemcpyHostToDevice, ch->ctx->stream));
    sync(ch->ctx);
    ch->D1[i] = D1;
    ch->D2[i] = D2;
}

void download_site(tn_chain* ch, int i, std::vector<double>& Mref)
{
    const int d = ch->d, D1 = ch->D1[i], D2 = ch->D2[i], D1p = pad32(D1), D2p = pad32(D2);
    std::vector<double> buf((size_t)D1p * d * D2p);
    TN_HIP(hipMemcpyAsync(buf.data(), ch->A[i], buf.size() * 8, hipMemcpyDeviceToHost, ch->ctx->stream));
    sync(ch->ctx);
    Mref.assign((size_t)d * D1 * D2, 0.0);
    for (int a = 0; a < D1; a++)
        for (int m = 0; m < d; m++)
            for (int b = 0; b < D2; b++) Mref[psi_ref(d, D1, m, a, b)] = buf[((size_t)a * d + m) * D2p + b];
}

// --- low precision environment copies -----------------------------------
void ensure_lo(tn_chain* ch, int prec)
{
    if (prec == TN_FP64) return;
    tn_chain::LoSet& s = ch->lo[prec];
    if (s.active) return;
    tn_ctx* c = ch->ctx;
    const size_t es = lowprec_bytes(prec);
    s.buf.ensure((size_t)2 * ch->L * ch->env_cap * es + 2 * 32 * 32 * ch->Dwmax * es + 1024);
    char* base = (char*)s.buf.p;
    s.L.assign(ch->L, nullptr);
    s.R.assign(ch->L, nullptr);
    for (int i = 0; i < ch->L; i++) {
        s.L[i] = base + (size_t)(2 * i) * ch->env_cap * es;
        s.R[i] = base + (size_t)(2 * i + 1) * ch->env_cap * es;
    }
    char* bb = base + (size_t)2 * ch->L * ch->env_cap * es;
    s.bL = bb;
    s.bR = bb + (size_t)32 * 32 * ch->Dwmax * es + 256;
    s.active = true;
    make_lowprec(c, prec, ch->bL, s.bL, (long)32 * 32 * ch->Dwmax);
    make_lowprec(c, prec, ch->bR, s.bR, (long)32 * 32 * ch->Dwmax);
    for (int i = 0; i < ch->L; i++) {
        make_lowprec(c, prec, ch->Lenv[i], s.L[i], (long)ch->env_cap, (long)ch->env_cap);
        make_lowprec(c, prec, ch->Renv[i], s.R[i], (long)ch->env_cap, (long)ch->env_cap);
    }
}

void drop_lo(tn_chain* ch)
{
    for (auto& s : ch->lo) {
        s.buf.release();
        s.active = false;
    }
}

// reduced-precision left/right environment feeding site i (boundary at the edges)
const void* lo_left(tn_chain* ch, int prec, int i)
{
    if (prec == TN_FP64) return nullptr;
    ensure_lo(ch, prec);
    return i > 0 ? ch->lo[prec].L[i - 1] : ch->lo[prec].bL;
}
const void* lo_right(tn_chain* ch, int prec, int i)
{
    if (prec == TN_FP64) return nullptr;
    ensure_lo(ch, prec);
    return i < ch->L - 1 ? ch->lo[prec].R[i + 1] : ch->lo[prec].bR;
}

// --- environment update of site i (dir 0: left env, dir 1: right env) ---
void env_update(tn_chain* ch, int i, int dir, const tn_sweep_opts& o, tn_sweep_stats* st)
{
    tn_ctx* c = ch->ctx;
    const int d = ch->d, D1p = pad32(ch->D1[i]), D2p = pad32(ch->D2[i]);
    const int Dw1 = ch->dw1[i], Dw2 = ch->dw2[i];
    const int eprec = o.env_precision;
    TN_REQUIRE(eprec >= TN_FP64 && eprec <= TN_BF16, TN_ERR_ARG,
               "env_precision must be FP64, FP32 or BF16 (the split-bfloat16 mode is for Heff)");
    double t0 = now_ms();
    if (dir == 0) {
        mix_build(ch->mixL[i], MIX_ENVL, ch->W[i].data(), Dw1, Dw2, d, D1p, D2p, c->stream);
        const double* Lin = i > 0 ? ch->Lenv[i - 1] : ch->bL;
        env_left(c, eprec, d, D1p, D2p, Dw1, Dw2, ch->mixL[i], Lin, lo_left(ch, eprec, i), ch->A[i], ch->Lenv[i],
                 o.gemm_tile);
        for (int p = TN_FP32; p <= TN_BF16X3; p++)
            if (ch->lo[p].active)
                make_lowprec(c, p, ch->Lenv[i], ch->lo[p].L[i], (long)D2p * Dw2 * D2p, (long)ch->env_cap);
    } else {
        mix_build(ch->mixR[i], MIX_ENVR, ch->W[i].data(), Dw1, Dw2, d, D1p, D2p, c->stream);
        const double* Rin = i < ch->L - 1 ? ch->Renv[i + 1] : ch->bR;
        env_right(c, eprec, d, D1p, D2p, Dw1, Dw2, ch->mixR[i], Rin, lo_right(ch, eprec, i), ch->A[i], ch->Renv[i],
                  o.gemm_tile);
        for (int p = TN_FP32; p <= TN_BF16X3; p++)
            if (ch->lo[p].active)
                make_lowprec(c, p, ch->Renv[i], ch->lo[p].R[i], (long)Dw1 * D1p * D1p, (long)ch->env_cap);
    }
    if (st) {
        sync(c);
        st->env_ms += now_ms() - t0;
        st->n_env++;
        const double D1 = ch->D1[i], D2 = ch->D2[i];
        if (dir == 0)
            st->env_flops += 2.0 * D1 * Dw1 * D1 * d * D2 + 2.0 * Dw1 * Dw2 * d * d * D1 * D2 + 2.0 * D2 * Dw2 * D2 * D1 * d;
        else
            st->env_flops += 2.0 * D1 * d * Dw2 * D2 * D2 + 2.0 * Dw1 * Dw2 * d * d * D1 * D2 + 2.0 * Dw1 * D1 * D1 * d * D2;
    }
}

// --- Heff apply callback for the Lanczos driver ---------------------------
void heff_timing_flush(tn_ctx* c)
{
    if (!c->heff_pending) return;
    heff_timing_read(c, &c->heff_pending->heff_gemm_ms, &c->heff_pending->heff_ms);
    c->heff_pending = nullptr;
}

void heff_cb(tn_ctx* c, const void* vop, const double* x, double* y)
{
    const HeffCall* hc = (const HeffCall*)vop;
    const bool timed = hc->st != nullptr && hc->timed;
    heff_timing_flush(c);  // the events are reused below
    heff_apply(c, hc->op, x, y, timed);
    if (timed) c->heff_pending = hc->st;
    if (hc->st) {
        hc->st->heff_flops += hc->op.flops_alg(hc->D1, hc->D2);
        hc->st->heff_gemm_flops += hc->op.flops_gemm(hc->D1, hc->D2);
        hc->st->n_matvec++;
        if (hc->op.precision == TN_FP64) hc->st->n_matvec_fp64++;
    }
}

// Identity channels of the environments of site i: for the first and the last
// channel of L (layout [D1p][Dw1][D1p]) and of R ([Dw2][D2p][D2p]), the largest
// deviation from the identity on the logical block (the padding is zero by
// construction).  A channel within tol of the identity is applied as one.
void env_identity(tn_chain* ch, int i, double tol, int* wl, int* wr)
{
    tn_ctx* c = ch->ctx;
    hipStream_t s = c->stream;
    const int D1 = ch->D1[i], D2 = ch->D2[i], D1p = pad32(D1), D2p = pad32(D2);
    const int Dw1 = ch->dw1[i], Dw2 = ch->dw2[i];
    *wl = *wr = -1;
    if (Dw1 < 2 && Dw2 < 2) return;
    const double* L = i > 0 ? ch->Lenv[i - 1] : ch->bL;
    const double* R = i < ch->L - 1 ? ch->Renv[i + 1] : ch->bR;
    c->ident.ensure(4 * 8);
    c->part.ensure(1024 * 8);
    double* dv = c->ident.as<double>();
    double* part = c->part.as<double>();
    const int chs[2] = {0, -1};
    for (int k = 0; k < 2; k++) {
        const int wlk = chs[k] < 0 ? Dw1 - 1 : 0, wrk = chs[k] < 0 ? Dw2 - 1 : 0;
        TN_HIP(eye_deviation(L + (long)wlk * D1p, (long)Dw1 * D1p, D1, dv + k, part, s));
        TN_HIP(eye_deviation(R + (long)wrk * D2p * D2p, D2p, D2, dv + 2 + k, part, s));
    }
    double h[4];
    d2h_small(c, h, dv, 4);
    auto ok = [&](double v) { return std::isfinite(v) && v <= tol; };
    if (Dw1 >= 2) *wl = ok(h[0]) ? 0 : (ok(h[1]) ? Dw1 - 1 : -1);
    if (Dw2 >= 2) *wr = ok(h[2]) ? 0 : (ok(h[3]) ? Dw2 - 1 : -1);
}

HeffOp make_heff(tn_chain* ch, int i, int prec, int tile, int wl, int wr, int mixop)
{
    HeffOp op;
    op.p = ch->d;
    op.D1p = pad32(ch->D1[i]);
    op.D2p = pad32(ch->D2[i]);
    op.Dw1 = ch->dw1[i];
    op.Dw2 = ch->dw2[i];
    op.L = i > 0 ? ch->Lenv[i - 1] : ch->bL;
    op.R = i < ch->L - 1 ? ch->Renv[i + 1] : ch->bR;
    op.precision = prec;
    op.Llo = lo_left(ch, prec, i);
    op.Rlo = lo_right(ch, prec, i);
    if (prec == TN_BF16X3) {
        op.Lps = i > 0 ? (long)ch->env_cap : (long)32 * 32 * ch->Dwmax;
        op.Rps = i < ch->L - 1 ? (long)ch->env_cap : (long)32 * 32 * ch->Dwmax;
    }
    op.wl_id = wl;
    op.wr_id = wr;
    mix_build(ch->mixH[i], mixop, ch->W[i].data(), op.Dw1, op.Dw2, ch->d, op.D1p, op.D2p, ch->ctx->stream,
              op.wl_id, op.wr_id);
    op.mix = &ch->mixH[i];
    op.tile = tile;
    return op;
}

// device copy of the diagonal blocks W[w][w'][m][m] of site i ([Dw1][Dw2][d])
const double* w_diag(tn_chain* ch, int i)
{
    tn_ctx* c = ch->ctx;
    const int d = ch->d, Dw1 = ch->dw1[i], Dw2 = ch->dw2[i];
    std::vector<double> wd((size_t)Dw1 * Dw2 * d);
    const double* W = ch->W[i].data();
    for (int w = 0; w < Dw1; w++)
        for (int v = 0; v < Dw2; v++)
            for (int m = 0; m < d; m++) wd[((size_t)w * Dw2 + v) * d + m] = W[(((size_t)w * Dw2 + v) * d + m) * d + m];
    c->dav_wd.ensure(wd.size() * 8);
    h2d(c, c->dav_wd.p, wd.data(), wd.size() * 8);
    return c->dav_wd.as<double>();
}

// unit roundoff of the Heff operands: the relative rounding of an apply
double unit_roundoff(int precision)
{
    return precision == TN_FP64 ? 0x1p-53 : (precision == TN_FP32 || precision == TN_BF16X3) ? 0x1p-24 : 0x1p-8;
}

// local eigensolve of site i; returns (e0, e1)
void eigensolve(tn_chain* ch, int i, const tn_sweep_opts& o, tn_sweep_stats* st, double* e0, double* e1)
{
    tn_ctx* c = ch->ctx;
    int wl = -1, wr = -1;
    if (o.skip_identity) env_identity(ch, i, ident_tol(o.env_precision), &wl, &wr);
    if (st) {
        st->n_ident_left += wl >= 0;
        st->n_ident_right += wr >= 0;
    }
    HeffCall hc;
    hc.op = make_heff(ch, i, o.precision, o.gemm_tile, wl, wr);
    hc.st = st;
    hc.timed = o.time_heff != 0;
    hc.D1 = ch->D1[i];
    hc.D2 = ch->D2[i];
    const long N = hc.op.n();
    double t0 = now_ms();
    // a reduced-precision operator cannot resolve residuals below its own
    // rounding (~2^-8 for BF16, ~2^-23 for FP32 operands, relative to |theta|)
    // when a refinement follows, stop there and let the refinement finish
    if (o.eig_method == 1) {
        TN_REQUIRE(o.precond >= 0 && o.precond <= 3, TN_ERR_ARG, "Davidson: precond must be 0..3");
        const double* Wd = w_diag(ch, i);
        // a learned correction gets one basis (no restart); when that does not
        // converge, plain Lanczos continues from its Ritz vector
        const bool learned = o.precond >= 2;
        const int cap = learned ? o.max_krylov : o.max_krylov * (1 + o.max_restarts);
        DavidsonResult dr = davidson(c, heff_cb, &hc, make_heff(ch, i, TN_FP64, o.gemm_tile, wl, wr), ch->D1[i],
                                     ch->D2[i], Wd, ch->A[i], o.eig_tol, o.max_krylov, cap, o.precond,
                                     o.precond_delta);
        *e0 = dr.e0;
        *e1 = dr.e1;
        bool conv = dr.converged;
        if (learned && !dr.converged) {
            if (st) st->n_precond_fallback++;
            LanczosResult r = lanczos(c, heff_cb, &hc, N, ch->A[i], o.eig_tol, o.max_krylov, o.max_restarts, 0.0,
                                      o.reorth == 1, unit_roundoff(TN_FP64), true);
            if (st) st->n_full_reorth += r.n_full;
            *e0 = r.e0;
            *e1 = r.e1;
            conv = r.converged;
        }
        if (st) (conv ? st->n_solve_converged : st->n_solve_capped)++;
        heff_timing_flush(c);
        if (st) st->lanczos_ms += now_ms() - t0;
        return;
    }
    const bool refining = o.refine && unit_roundoff(o.refine_precision) < unit_roundoff(o.precision);
    const double rel = refining ? (o.precision == TN_BF16 ? 4e-3 : 4e-7) : 0.0;
    LanczosResult r = lanczos(c, heff_cb, &hc, N, ch->A[i], o.eig_tol, o.max_krylov, o.max_restarts, rel,
                              o.reorth == 1, unit_roundoff(o.precision), true);
    if (st) st->n_full_reorth += r.n_full;
    if (refining) {
        // polish the reduced-precision Ritz vector with a more accurate operator
        HeffCall hr = hc;
        hr.op = make_heff(ch, i, o.refine_precision, o.gemm_tile, wl, wr);
        LanczosResult r2 = lanczos(c, heff_cb, &hr, N, ch->A[i], o.refine_tol, o.max_krylov, o.max_restarts, 0.0,
                                   o.reorth == 1, unit_roundoff(o.refine_precision), true);
        if (st) st->n_full_reorth += r2.n_full;
        r.e0 = r2.e0;
        r.e1 = r2.e1;
        r.converged = r2.converged;
    }
    if (st) (r.converged ? st->n_solve_converged : st->n_solve_capped)++;
    heff_timing_flush(c);
    if (st) st->lanczos_ms += now_ms() - t0;
    *e0 = r.e0;
    *e1 = r.e1;
}

// Bond.C:86-106 SetSingularValues
void bond_set_sv(tn_chain* ch, int b, const std::vector<double>& s, double comp)
{
    const int n = (int)s.size();
    double nr = 0;
    for (double v : s) nr += v * v;
    nr = std::sqrt(nr);
    std::vector<double> sn(n), old(n, 0.0);
    for (int k = 0; k < n; k++) sn[k] = nr > 0 ? s[k] / nr : 0.0;
    for (int k = 0; k < n && k < (int)ch->lam[b].size(); k++) old[k] = ch->lam[b][k];
    double md = 0;
    for (int k = 0; k < n; k++) md = std::max(md, std::fabs(sn[k] - old[k]));
    ch->maxDelta[b] = md;
    ch->lam[b] = sn;
    ch->compErr[b] += comp;
}

// resize a site keeping the overlapping block (MPSSite.C:156-171), host path
void site_resize(tn_chain* ch, int i, int nD1, int nD2, bool corner_rule)
{
    const int d = ch->d, oD1 = ch->D1[i], oD2 = ch->D2[i];
    if (oD1 == nD1 && oD2 == nD2) return;
    std::vector<double> M;
    download_site(ch, i, M);
    std::vector<double> N((size_t)d * nD1 * nD2, 0.0);
    for (int m = 0; m < d; m++)
        for (int a = 0; a < std::min(oD1, nD1); a++)
            for (int b = 0; b < std::min(oD2, nD2); b++) N[psi_ref(d, nD1, m, a, b)] = M[psi_ref(d, oD1, m, a, b)];
    if (corner_rule)
        for (int m = 0; m < d; m++)
            for (int a = oD1 - 1; a < nD1; a++)
                for (int b = oD2 - 1; b < nD2; b++)
                    if (a >= 0 && b >= 0) N[psi_ref(d, nD1, m, a, b)] = 0.0;
    upload_site(ch, i, nD1, nD2, N.data());
}

void rescale_site(tn_chain* ch, int i)
{
    tn_ctx* c = ch->ctx;
    const long n = (long)pad32(ch->D1[i]) * ch->d * pad32(ch->D2[i]);
    c->lanczos_small.ensure(sizeof(double) * (4 * 512 + 64));
    double* h = c->lanczos_small.as<double>() + 4 * 512 + 16;
    long pcap = mdot_part_size(n, 1);
    c->part.ensure(pcap * 8);
    TN_HIP(mdot(ch->A[i], n, 1, ch->A[i], n, h, c->part.as<double>(), (int)pcap, c->stream));
    TN_HIP(dsqrt_scalar(h, h, c->stream));
    TN_HIP(scale_by_inv(ch->A[i], h, n, c->stream));
}

// The neighbour receives s/|s| rather than s: the overall norm of the state
// is irrelevant to the sweep (the site eigenproblem and the edge rescale
// normalise it), and forwarding raw s multiplies the tensors of a random
// high-D start by ~0.02 per site, which underflows FP64 along a long chain.
void normalised_copy(const std::vector<double>& s, std::vector<double>& out)
{
    double nr = 0;
    for (double v : s) nr += v * v;
    nr = std::sqrt(nr);
    for (size_t k = 0; k < s.size(); k++) out[k] = nr > 0 ? s[k] / nr : 0.0;
}

// MPSSite_Normalization.C:80-107 NormalizeQR: A = Q R (left move) or A = L Q
// (right move) by shifted CholeskyQR; the bond's singular values are cleared
// (Bond::TransferQR -> ClearSingularValues, Bond.C:108-135).  Returns false
// (nothing changed) when the matrix is wide or CholeskyQR is not accurate.
bool qr_normalize(tn_chain* ch, int i, int dir, tn_sweep_stats* st)
{
    int shifted = 0;
    tn_ctx* c = ch->ctx;
    hipStream_t s = c->stream;
    const int d = ch->d;
    const int D1 = ch->D1[i], D2 = ch->D2[i], D1p = pad32(D1), D2p = pad32(D2);
    const double tol = 1e-11;
    if (dir == 0) {
        if (D1 * d < D2) return false;
        const int m = D1p * d, n = D2p;
        c->spare.ensure((size_t)n * n * 8);
        double* Lacc = c->spare.as<double>();
        if (!cholqr_device(c, ch->A[i], D2p, m, n, D2, ch->spare, Lacc, tol, &shifted, false)) return false;
        if (st) st->n_qr_shifted += shifted;
        // site i <- Q (already in ch->spare: swap buffers)
        std::swap(ch->A[i], ch->spare);
        // neighbour: M_{i+1} <- R M_{i+1},  R = Lacc^T
        const int j = i + 1, D3p = pad32(ch->D2[j]);
        GemmDesc g{};
        g.tin = T_F64; g.tout = T_F64; g.a_kmajor = false; g.b_kmajor = false;
        g.M = n; g.N = d * D3p; g.K = n; g.batch = 1;
        g.A = Lacc; g.lda = n; g.B = ch->A[j]; g.ldb = (long)d * D3p;
        g.C = ch->spare; g.ldc = (long)d * D3p; g.alpha = 1.0f;
        gemm_run(c, g);
        std::swap(ch->A[j], ch->spare);
        rescale_site(ch, j);  // keep the norm O(1) along the chain (see normalised_copy)
        ch->lam[i].assign(D2, 0.0);
        ch->maxDelta[i] = 0.0;
        return true;
    }
    if (d * D2 < D1) return false;
    // A^T ((d*D2p) x D1p) by transposed copy, then A^T = Q R  ->  A = R^T Q^T
    const int m = d * D2p, n = D1p;
    c->svd_A.ensure((size_t)m * n * 8);
    double* At = c->svd_A.as<double>();
    TN_HIP(copy2(ch->A[i], (long)d * D2p, 1, D1p, d * D2p, At, 1, D1p, D1p, d * D2p, s));
    c->spare.ensure((size_t)n * n * 8);
    double* Lacc = c->spare.as<double>();
    c->svd_U.ensure((size_t)m * n * 8);
    double* Qt = c->svd_U.as<double>();
    if (!cholqr_device(c, At, n, m, n, D1, Qt, Lacc, tol, &shifted, false)) return false;
    if (st) st->n_qr_shifted += shifted;
    // site i <- Q^T
    TN_HIP(copy2(Qt, 1, n, D1p, d * D2p, ch->A[i], (long)d * D2p, 1, D1p, d * D2p, s));
    // neighbour: M_{i-1} <- M_{i-1} Lacc
    const int j = i - 1, D0p = pad32(ch->D1[j]);
    GemmDesc g{};
    g.tin = T_F64; g.tout = T_F64; g.a_kmajor = true; g.b_kmajor = false;
    g.M = D0p * d; g.N = n; g.K = n; g.batch = 1;
    g.A = ch->A[j]; g.lda = n; g.B = Lacc; g.ldb = n;
    g.C = ch->spare; g.ldc = n; g.alpha = 1.0f;
    // Lacc lives in ctx->spare, the product goes to the chain's spare buffer
    gemm_run(c, g);
    std::swap(ch->A[j], ch->spare);
    rescale_site(ch, j);
    ch->lam[i - 1].assign(D1, 0.0);
    ch->maxDelta[i - 1] = 0.0;
    return true;
}

// MPSSite_Normalization.C:45-78 SVDNormalize(lr, nullptr) incl. HandleEdges :20-43
void svd_normalize(tn_chain* ch, int i, int dir, tn_sweep_stats* st, int gauge = 0)
{
    tn_ctx* c = ch->ctx;
    hipStream_t s = c->stream;
    const int d = ch->d, L = ch->L;
    const int D1 = ch->D1[i], D2 = ch->D2[i], D1p = pad32(D1), D2p = pad32(D2);
    if (dir == 1 && i == 0) {
        const int newD2 = L > 1 ? std::max((int)ch->lam[0].size(), d) : D2;
        if (newD2 < D2) site_resize(ch, i, D1, newD2, false);
        rescale_site(ch, i);
        return;
    }
    if (dir == 0 && i == L - 1) {
        const int newD1 = L > 1 ? std::max((int)ch->lam[L - 2].size(), d) : D1;
        if (newD1 < D1) site_resize(ch, i, newD1, D2, false);
        rescale_site(ch, i);
        return;
    }
    double t0 = now_ms();
    if (gauge == 1 && qr_normalize(ch, i, dir, st)) {
        if (st) {
            sync(c);
            st->svd_ms += now_ms() - t0;
            st->n_qr++;
        }
        return;
    }
    if (gauge == 1 && st) st->n_qr_fallback++;
    std::vector<double> sv;
    if (dir == 0) {
        // A as (D1*d) x D2: row (a,m) at (a*d+m)*D2p, column b
        const int r = D1 * d, cc = D2, k = std::min(r, cc), kp = pad32(k);
        c->svd_U.ensure((size_t)r * k * 8 + 64);
        c->svd_Vt.ensure((size_t)kp * D2p * 8 + 64);
        svd_device(c, ch->A[i], D2p, 1, r, cc, c->svd_U.as<double>(), k, r, c->svd_Vt.as<double>(), D2p, kp, sv);
        bond_set_sv(ch, i, sv, 0.0);
        // neighbour: M_{i+1} <- diag(s) Vt M_{i+1}
        const int j = i + 1, D3p = pad32(ch->D2[j]);
        std::vector<double> sv_pad(kp, 0.0);
        normalised_copy(sv, sv_pad);
        c->pack.ensure(kp * 8 + 64);
        h2d(c, c->pack.p, sv_pad.data(), kp * 8);
        TN_HIP(scale_rows(c->svd_Vt.as<double>(), D2p, kp, D2p, c->pack.as<double>(), s));
        GemmDesc g{};
        g.tin = T_F64; g.tout = T_F64; g.a_kmajor = true; g.b_kmajor = false;
        g.M = kp; g.N = d * D3p; g.K = D2p; g.batch = 1;
        g.A = c->svd_Vt.p; g.lda = D2p;
        g.B = ch->A[j]; g.ldb = (long)d * D3p;
        g.C = ch->spare; g.ldc = (long)d * D3p; g.alpha = 1.0f;
        gemm_run(c, g);
        std::swap(ch->A[j], ch->spare);
        ch->D1[j] = k;
        // site i <- U  (rows D1p*d, cols kp, zero padded)
        TN_HIP(copy2(c->svd_U.as<double>(), k, 1, r, k, ch->A[i], kp, 1, D1p * d, kp, s));
        ch->D2[i] = k;
    } else {
        // A as D1 x (d*D2): pack the logical matrix first
        const int r = D1, cc = d * D2, k = std::min(r, cc), kp = pad32(k);
        c->svd_A.ensure((size_t)r * cc * 8 + 64);
        TN_HIP(copy3(ch->A[i], (long)d * D2p, D2p, 1, c->svd_A.as<double>(), r, d, D2, D2, s));
        c->svd_U.ensure((size_t)D1p * kp * 8 + 64);
        c->svd_Vt.ensure((size_t)kp * cc * 8 + 64);
        svd_device(c, c->svd_A.as<double>(), cc, 1, r, cc, c->svd_U.as<double>(), kp, D1p,
                   c->svd_Vt.as<double>(), cc, kp, sv);
        bond_set_sv(ch, i - 1, sv, 0.0);
        // neighbour: M_{i-1} <- M_{i-1} U diag(s)
        const int j = i - 1, D0p = pad32(ch->D1[j]);
        std::vector<double> sv_pad(kp, 0.0);
        normalised_copy(sv, sv_pad);
        c->pack.ensure(kp * 8 + 64);
        h2d(c, c->pack.p, sv_pad.data(), kp * 8);
        TN_HIP(scale_cols(c->svd_U.as<double>(), kp, D1p, kp, c->pack.as<double>(), s));
        GemmDesc g{};
        g.tin = T_F64; g.tout = T_F64; g.a_kmajor = true; g.b_kmajor = false;
        g.M = D0p * d; g.N = kp; g.K = D1p; g.batch = 1;
        g.A = ch->A[j]; g.lda = D1p;
        g.B = c->svd_U.p; g.ldb = kp;
        g.C = ch->spare; g.ldc = kp; g.alpha = 1.0f;
        gemm_run(c, g);
        std::swap(ch->A[j], ch->spare);
        ch->D2[j] = k;
        // site i <- Vt (k x d*D2) into [kp][d][D2p]
        TN_HIP(copy3(c->svd_Vt.as<double>(), cc, D2, 1, ch->A[i], kp, d, D2, D2p, s));
        ch->D1[i] = k;
    }
    if (st) {
        sync(c);
        st->svd_ms += now_ms() - t0;
        st->n_svd++;
    }
}

void refine_site(tn_chain* ch, int i, int dir, const tn_sweep_opts& o, tn_sweep_stats* st)
{
    double e0 = 0, e1 = 0;
    eigensolve(ch, i, o, st, &e0, &e1);
    ch->IterDE[i] = e0 - ch->Emin[i];
    ch->Emin[i] = e0;
    ch->GapE[i] = e1 - 1.0;  // MPSSite_Variational.C:29 as written
    if (st) st->energy = e0;
    svd_normalize(ch, i, dir, st, o.gauge);
    env_update(ch, i, dir, o, st);
}

// mixing-table and Lanczos helpers for arbitrary physical dimension p
void two_site_W(int d, int Dw1, int Dwm, int Dw3, const double* Wa, const double* Wb, std::vector<double>& W12)
{
    const int p = d * d;
    W12.assign((size_t)Dw1 * Dw3 * p * p, 0.0);
    for (int w = 0; w < Dw1; w++)
        for (int w2 = 0; w2 < Dw3; w2++)
            for (int wp = 0; wp < Dwm; wp++)
                for (int m1 = 0; m1 < d; m1++)
                    for (int n1 = 0; n1 < d; n1++) {
                        double a = Wa[((w * Dwm + wp) * d + m1) * d + n1];
                        if (a == 0.0) continue;
                        for (int m2 = 0; m2 < d; m2++)
                            for (int n2 = 0; n2 < d; n2++)
                                W12[(((size_t)w * Dw3 + w2) * p + (m1 * d + m2)) * p + (n1 * d + n2)] +=
                                    a * Wb[((wp * Dw3 + w2) * d + m2) * d + n2];
                    }
}

// grow the per-site device buffers so every bond dimension up to Dcap fits;
// site data are kept, environments must be reloaded afterwards
void chain_reserve(tn_chain* ch, int Dcap)
{
    if (Dcap <= ch->Dcap) return;
    tn_ctx* c = ch->ctx;
    sync(c);
    const int Dcapp = pad32(Dcap);
    const size_t site_cap = (size_t)Dcapp * ch->d * Dcapp;
    const size_t env_cap = (size_t)Dcapp * ch->Dwmax * Dcapp;
    DBuf nsites;
    nsites.ensure(site_cap * 8 * ch->L);
    TN_HIP(fill_zero(nsites.p, (long)nsites.bytes, c->stream));
    for (int i = 0; i < ch->L; i++) {
        const size_t n = (size_t)pad32(ch->D1[i]) * ch->d * pad32(ch->D2[i]);
        TN_HIP(hipMemcpyAsync(nsites.as<double>() + (size_t)i * site_cap, ch->A[i], n * 8,
                              hipMemcpyDeviceToDevice, c->stream));
    }
    sync(c);
    ch->sites_buf.release();
    ch->sites_buf = nsites;
    ch->env_buf.release();
    ch->spare_buf.release();
    drop_lo(ch);
    ch->Dcap = Dcap;
    ch->Dcapp = Dcapp;
    ch->site_cap = site_cap;
    ch->env_cap = env_cap;
    ch->spare_buf.ensure(site_cap * 8);
    ch->env_buf.ensure(env_cap * 8 * 2 * ch->L);
    TN_HIP(fill_zero(ch->env_buf.p, (long)ch->env_buf.bytes, c->stream));
    for (int i = 0; i < ch->L; i++) {
        ch->A[i] = ch->sites_buf.as<double>() + (size_t)i * site_cap;
        ch->Lenv[i] = ch->env_buf.as<double>() + (size_t)(2 * i) * env_cap;
        ch->Renv[i] = ch->env_buf.as<double>() + (size_t)(2 * i + 1) * env_cap;
    }
    ch->spare = ch->spare_buf.as<double>();
    sync(c);
}

}  // namespace tn

// ---------------------------------------------------------------------------
// Work space of the allocation-free Heff apply (tn_heff_apply_device_noalloc):
// one file-scope device array, which every GPU the library runs on holds until
// the process ends.  Host code never uses the array's name as an address: its
// device address is looked up per device (hipGetSymbolAddress, with that
// device current) when a context is created, and only that address is passed
// to kernels.
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// device address of this device's work space array (the device must be current)
static void* resolve_scratch(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    TN_REQUIRE(device >= 0 && device < kMaxDevices, TN_ERR_ARG, "device index beyond the work space table");
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        TN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TN_REQUIRE(p != nullptr, TN_ERR_HIP, "hipGetSymbolAddress returned no address for the work space");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// ===========================================================================
// C ABI
extern "C" {

int tn_version(void) { return 1; }

int tn_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int tn_ctx_create(int device, tn_ctx** out)
{
    if (!out) return TN_ERR_ARG;
    *out = nullptr;
    tn_ctx* c = new tn_ctx();
    ABI_BEGIN
    int n = 0;
    TN_HIP(hipGetDeviceCount(&n));
    TN_REQUIRE(device >= 0 && device < n, TN_ERR_ARG, "no such HIP device");
    c->device = device;
    TN_HIP(hipSetDevice(device));
    TN_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) TN_HIP(hipEventCreate(&e));
    const size_t pub = (size_t)TN_PUBLISH_MAX * sizeof(double);
    TN_HIP(hipHostMalloc((void**)&c->sc.h, pub + 256, hipHostMallocCoherent | hipHostMallocMapped));
    TN_HIP(hipHostGetDevicePointer((void**)&c->sc.d, c->sc.h, 0));
    c->sc.flag_h = (unsigned long long*)((char*)c->sc.h + pub);
    c->sc.flag_d = (unsigned long long*)((char*)c->sc.d + pub);
    *c->sc.flag_h = 0;
    c->noalloc = resolve_scratch(device);
    *out = c;
    return TN_OK;
    }
    catch (const TnError& e) {
        int code = e.code;
        delete c;
        return code;
    }
}

int tn_ctx_set_autotune(tn_ctx* c, int on)
{
    if (!c) return TN_ERR_ARG;
    c->autotune = on != 0;
    return TN_OK;
}

int tn_ctx_tune_report(tn_ctx* c, char* buf, int len)
{
    if (!c || !buf || len <= 0) return TN_ERR_ARG;
    std::string out;
    static const char* tn_[] = {"f64", "f32", "bf16", "bf16x3"};
    for (const auto& kv : c->tune) {
        const auto& k = kv.first;
        char line[192];
        snprintf(line, sizeof(line), "%s->%s %c%c M=%ld N=%ld K=%ld batch=%ld acc=%ld : tile %s ksplit %d\n",
                 tn_[k[0]], tn_[k[1]], k[2] ? 'K' : 'T', k[3] ? 'K' : 'T', k[4], k[5], k[6], k[7], k[8],
                 gemm_tile_name(kv.second.first), kv.second.second);
        out += line;
    }
    snprintf(buf, (size_t)len, "%s", out.c_str());
    return (int)c->tune.size();
}

int tn_ctx_tune_load(tn_ctx* c, const char* path)
{
    if (!c || !path) return TN_ERR_ARG;
    FILE* f = fopen(path, "r");
    if (!f) {
        c->err = std::string("tn_ctx_tune_load: cannot open ") + path;
        return TN_ERR_ARG;
    }
    int n = 0;
    long v[11];
    while (fscanf(f, "%ld %ld %ld %ld %ld %ld %ld %ld %ld %ld %ld", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6],
                  &v[7], &v[8], &v[9], &v[10]) == 11) {
        // a malformed line is skipped: element types in {f64, f32, bf16},
        // layout and accumulate flags 0/1, positive shape, a known tile and k-split
        const bool types_ok = v[0] >= 0 && v[0] <= 3 && v[1] >= 0 && v[1] <= 2;
        const bool flags_ok = (v[2] == 0 || v[2] == 1) && (v[3] == 0 || v[3] == 1) && (v[8] == 0 || v[8] == 1);
        const bool shape_ok = v[4] > 0 && v[5] > 0 && v[6] > 0 && v[7] > 0;
        if (!types_ok || !flags_ok || !shape_ok) continue;
        if (v[9] < 1 || v[9] > gemm_num_tiles() || (v[10] != 1 && v[10] != 2 && v[10] != 4)) continue;
        if (!gemm_tile_ok((int)v[9], (int)v[0], (int)v[4], (int)v[5], (int)v[6])) continue;
        const std::array<long, 9> key{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]};
        c->tune[key] = {(int)v[9], (int)v[10]};
        n++;
    }
    fclose(f);
    return n;
}

int tn_ctx_tune_save(tn_ctx* c, const char* path)
{
    if (!c || !path) return TN_ERR_ARG;
    FILE* f = fopen(path, "w");
    if (!f) {
        c->err = std::string("tn_ctx_tune_save: cannot open ") + path;
        return TN_ERR_ARG;
    }
    for (const auto& kv : c->tune) {
        const auto& k = kv.first;
        fprintf(f, "%ld %ld %ld %ld %ld %ld %ld %ld %ld %d %d\n", k[0], k[1], k[2], k[3], k[4], k[5], k[6], k[7], k[8],
                kv.second.first, kv.second.second);
    }
    fclose(f);
    return (int)c->tune.size();
}

int tn_ctx_destroy(tn_ctx* c)
{
    if (!c) return TN_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    tn::DBuf* bufs[] = {&c->T1, &c->T2, &c->xlo, &c->part, &c->krylov, &c->lanczos_small, &c->svd_X,
                        &c->svd_V, &c->svd_U, &c->svd_Vt, &c->svd_s, &c->svd_perm, &c->svd_pairs,
                        &c->svd_flag, &c->svd_A, &c->spare, &c->pack, &c->qr_W, &c->qr_G, &c->qr_L,
                        &c->qr_T, &c->qr_linv, &c->qr_Q, &c->tune_C, &c->splitk, &c->b_L, &c->b_R,
                        &c->b_Llo, &c->b_Rlo, &c->b_x, &c->b_y, &c->b_M, &c->dav_W, &c->dav_vec,
                        &c->dav_wd, &c->mlp, &c->ident, &c->defl, &c->svd_R, &c->svd_Q, &c->svd_H,
                        &c->svd_W, &c->svd_X2};
    for (auto* b : bufs) b->release();
    c->b_mix_heff.buf.release();
    c->b_mix_envl.buf.release();
    c->b_mix_envr.buf.release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->sc.h) (void)hipHostFree(c->sc.h);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TN_OK;
}

const char* tn_last_error(const tn_ctx* c) { return c ? c->err.c_str() : "null context"; }

void tn_sweep_opts_default(tn_sweep_opts* o)
{
    if (!o) return;
    o->eig_tol = 1e-10;
    o->max_krylov = 64;
    o->max_restarts = 8;
    o->precision = TN_FP64;
    o->env_precision = TN_FP64;
    o->refine = 1;
    o->refine_tol = 1e-10;
    o->gemm_tile = 0;
    o->gauge = 0;
    o->refine_precision = TN_FP64;
    o->eig_method = 0;
    o->precond = 1;
    o->precond_delta = 1e-2;
    o->skip_identity = 1;
    o->time_heff = 1;
    o->reorth = 0;
}

int tn_svcompress_rank(int N, const double* s, int Dmax, double eps, double* discarded, double* rescale)
{
    // SVCompressorImp.C:19-73
    if (N < 0 || (N > 0 && !s)) return TN_ERR_ARG;
    int D = N;
    double integrated = 0.0;
    if (eps == 0.0) {
        D = Dmax;
        for (int id = D; id < N; id++) integrated += s[id] * s[id];
    } else {
        for (; D >= 1; D--) {
            double s2 = s[D - 1] * s[D - 1];
            if (integrated + s2 >= eps) break;
            integrated += s2;
        }
        if (Dmax > 0 && D > Dmax) D = Dmax;
    }
    double r = 1.0;
    if (D < N) {
        double before = 0, after = 0;
        for (int i = 0; i < N; i++) before += s[i] * s[i];
        for (int i = 0; i < D; i++) after += s[i] * s[i];
        if (after > 0.0) r = before / after;
    }
    if (D > N) D = N;
    if (discarded) *discarded = integrated;
    if (rescale) *rescale = r;
    return D;
}

// ---- seam primitives ------------------------------------------------------
int tn_site_bind(tn_ctx* c, int d, int D1, int D2, int Dw1, int Dw2, const double* W, const double* L,
                 const double* R, int precision)
{
    if (!c) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(d > 0 && D1 > 0 && D2 > 0 && Dw1 > 0 && Dw2 > 0 && W && L && R, TN_ERR_ARG, "tn_site_bind: bad arguments");
    TN_REQUIRE(precision >= TN_FP64 && precision <= TN_BF16X3, TN_ERR_ARG, "tn_site_bind: bad precision");
    const int D1p = pad32(D1), D2p = pad32(D2);
    c->b_d = d; c->b_D1 = D1; c->b_D2 = D2; c->b_Dw1 = Dw1; c->b_Dw2 = Dw2; c->b_prec = precision;
    c->b_W.assign(W, W + (size_t)Dw1 * Dw2 * d * d);
    // L(w,i,j) = L[w + Dw1*(i + D1*j)] -> Lint[i][w][j] padded
    std::vector<double> Li((size_t)D1p * Dw1 * D1p, 0.0), Ri((size_t)Dw2 * D2p * D2p, 0.0);
    for (int i = 0; i < D1; i++)
        for (int w = 0; w < Dw1; w++)
            for (int j = 0; j < D1; j++) Li[((size_t)i * Dw1 + w) * D1p + j] = L[w + (size_t)Dw1 * (i + (size_t)D1 * j)];
    for (int w = 0; w < Dw2; w++)
        for (int i = 0; i < D2; i++)
            for (int j = 0; j < D2; j++) Ri[((size_t)w * D2p + i) * D2p + j] = R[w + (size_t)Dw2 * (i + (size_t)D2 * j)];
    c->b_L.ensure(Li.size() * 8);
    c->b_R.ensure(Ri.size() * 8);
    h2d(c, c->b_L.p, Li.data(), Li.size() * 8);
    h2d(c, c->b_R.p, Ri.data(), Ri.size() * 8);
    if (precision != TN_FP64) {
        c->b_Llo.ensure(Li.size() * lowprec_bytes(precision));
        c->b_Rlo.ensure(Ri.size() * lowprec_bytes(precision));
        make_lowprec(c, precision, c->b_L.as<double>(), c->b_Llo.p, (long)Li.size());
        make_lowprec(c, precision, c->b_R.as<double>(), c->b_Rlo.p, (long)Ri.size());
    }
    // first or last channel equal to the identity (FP64 data: within 1e-12)
    auto dev_from_eye = [](const std::vector<double>& E, long sw, long sr, int w, int D) {
        double m = 0.0;
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++) {
                const double v = E[w * sw + i * sr + j] - (i == j ? 1.0 : 0.0);
                m = std::isfinite(v) ? std::max(m, std::fabs(v)) : INFINITY;
            }
        return m;
    };
    auto pick = [&](const std::vector<double>& E, long sw, long sr, int Dw, int D) {
        if (Dw < 2) return -1;
        if (dev_from_eye(E, sw, sr, 0, D) <= kIdentTolFp64) return 0;
        if (dev_from_eye(E, sw, sr, Dw - 1, D) <= kIdentTolFp64) return Dw - 1;
        return -1;
    };
    c->b_wl = pick(Li, D1p, (long)Dw1 * D1p, Dw1, D1);
    c->b_wr = pick(Ri, (long)D2p * D2p, D2p, Dw2, D2);
    mix_build(c->b_mix_heff, MIX_HEFF_T, c->b_W.data(), Dw1, Dw2, d, D1p, D2p, c->stream, c->b_wl, c->b_wr);
    const long n = (long)D1p * d * D2p;
    c->b_x.ensure(n * 8);
    c->b_y.ensure(n * 8);
    sync(c);
    c->bound = true;
    return TN_OK;
    ABI_END(c)
}

static HeffOp bound_op(tn_ctx* c)
{
    HeffOp op;
    op.p = c->b_d;
    op.D1p = pad32(c->b_D1);
    op.D2p = pad32(c->b_D2);
    op.Dw1 = c->b_Dw1;
    op.Dw2 = c->b_Dw2;
    op.L = c->b_L.as<double>();
    op.R = c->b_R.as<double>();
    op.Llo = c->b_prec == TN_FP64 ? nullptr : c->b_Llo.p;
    op.Rlo = c->b_prec == TN_FP64 ? nullptr : c->b_Rlo.p;
    if (c->b_prec == TN_BF16X3) {
        op.Lps = (long)pad32(c->b_D1) * c->b_Dw1 * pad32(c->b_D1);
        op.Rps = (long)c->b_Dw2 * pad32(c->b_D2) * pad32(c->b_D2);
    }
    op.precision = c->b_prec;
    op.mix = &c->b_mix_heff;
    op.tile = 0;
    op.wl_id = c->b_wl;
    op.wr_id = c->b_wr;
    return op;
}

static void psi_to_dev(tn_ctx* c, const double* x, double* dst)
{
    const int d = c->b_d, D1 = c->b_D1, D2 = c->b_D2, D2p = pad32(D2), D1p = pad32(D1);
    std::vector<double> buf((size_t)D1p * d * D2p, 0.0);
    for (int a = 0; a < D1; a++)
        for (int m = 0; m < d; m++)
            for (int b = 0; b < D2; b++) buf[((size_t)a * d + m) * D2p + b] = x[psi_ref(d, D1, m, a, b)];
    h2d(c, dst, buf.data(), buf.size() * 8);
}
static void psi_from_dev(tn_ctx* c, const double* src, double* y)
{
    const int d = c->b_d, D1 = c->b_D1, D2 = c->b_D2, D2p = pad32(D2), D1p = pad32(D1);
    std::vector<double> buf((size_t)D1p * d * D2p);
    d2h(c, buf.data(), src, buf.size() * 8);
    for (int a = 0; a < D1; a++)
        for (int m = 0; m < d; m++)
            for (int b = 0; b < D2; b++) y[psi_ref(d, D1, m, a, b)] = buf[((size_t)a * d + m) * D2p + b];
}

int tn_heff_apply(tn_ctx* c, const double* x, double* y, int64_t N)
{
    if (!c) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(c->bound, TN_ERR_STATE, "tn_heff_apply: no site bound");
    TN_REQUIRE(x && y && N == (int64_t)c->b_d * c->b_D1 * c->b_D2, TN_ERR_ARG, "tn_heff_apply: N != d*D1*D2");
    psi_to_dev(c, x, c->b_x.as<double>());
    HeffOp op = bound_op(c);
    heff_apply(c, op, c->b_x.as<double>(), c->b_y.as<double>(), false);
    sync(c);
    psi_from_dev(c, c->b_y.as<double>(), y);
    return TN_OK;
    ABI_END(c)
}

int tn_env_update(tn_ctx* c, int dir, const double* M, double* Fout)
{
    if (!c) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(c->bound, TN_ERR_STATE, "tn_env_update: no site bound");
    TN_REQUIRE(M && Fout && (dir == 0 || dir == 1), TN_ERR_ARG, "tn_env_update: bad arguments");
    const int d = c->b_d, D1 = c->b_D1, D2 = c->b_D2, D1p = pad32(D1), D2p = pad32(D2);
    const int Dw1 = c->b_Dw1, Dw2 = c->b_Dw2;
    c->b_M.ensure((size_t)D1p * d * D2p * 8);
    psi_to_dev(c, M, c->b_M.as<double>());
    // the split-bfloat16 binding updates environments in FP64
    const int prec = c->b_prec == TN_BF16X3 ? TN_FP64 : c->b_prec;
    if (dir == 0) {
        mix_build(c->b_mix_envl, MIX_ENVL, c->b_W.data(), Dw1, Dw2, d, D1p, D2p, c->stream);
        c->b_y.ensure((size_t)D2p * Dw2 * D2p * 8);
        env_left(c, prec, d, D1p, D2p, Dw1, Dw2, c->b_mix_envl, c->b_L.as<double>(), c->b_Llo.p,
                 c->b_M.as<double>(), c->b_y.as<double>(), 0);
        sync(c);
        std::vector<double> buf((size_t)D2p * Dw2 * D2p);
        d2h(c, buf.data(), c->b_y.p, buf.size() * 8);
        for (int i = 0; i < D2; i++)
            for (int w = 0; w < Dw2; w++)
                for (int j = 0; j < D2; j++) Fout[w + (size_t)Dw2 * (i + (size_t)D2 * j)] = buf[((size_t)i * Dw2 + w) * D2p + j];
    } else {
        mix_build(c->b_mix_envr, MIX_ENVR, c->b_W.data(), Dw1, Dw2, d, D1p, D2p, c->stream);
        c->b_y.ensure((size_t)Dw1 * D1p * D1p * 8);
        env_right(c, prec, d, D1p, D2p, Dw1, Dw2, c->b_mix_envr, c->b_R.as<double>(), c->b_Rlo.p,
                  c->b_M.as<double>(), c->b_y.as<double>(), 0);
        sync(c);
        std::vector<double> buf((size_t)Dw1 * D1p * D1p);
        d2h(c, buf.data(), c->b_y.p, buf.size() * 8);
        for (int w = 0; w < Dw1; w++)
            for (int i = 0; i < D1; i++)
                for (int j = 0; j < D1; j++) Fout[w + (size_t)Dw1 * (i + (size_t)D1 * j)] = buf[((size_t)w * D1p + i) * D1p + j];
    }
    // restore the y buffer size for later applies
    c->b_y.ensure((size_t)D1p * d * D2p * 8);
    return TN_OK;
    ABI_END(c)
}

static void bound_cb(tn_ctx* c, const void* vop, const double* x, double* y)
{
    heff_apply(c, *(const HeffOp*)vop, x, y, false);
}

// H' = H + sigma x0 x0^T: the lowest eigenpair of H' is the second of H when
// sigma exceeds lambda_2 - lambda_1 (x0 the lowest eigenvector, unit norm)
struct DeflatedOp {
    const HeffOp* op;
    const double* x0;
    double sigma;
    double* dot;
    double* part;
    int pcap;
    long N;
};
static void deflated_cb(tn_ctx* c, const void* vop, const double* x, double* y)
{
    const DeflatedOp* d = (const DeflatedOp*)vop;
    heff_apply(c, *d->op, x, y, false);
    TN_HIP(mdot(d->x0, d->N, 1, x, d->N, d->dot, d->part, d->pcap, c->stream));
    TN_HIP(gemv_acc(d->x0, d->N, 1, d->dot, d->sigma, 1.0, y, d->N, c->stream));
}

// seeded U(-1,1) start vector over the logical entries (reference layout);
// the padding of the internal layout stays zero, so the operator's null
// space on the padding never enters the Krylov space
static void random_start(double* v, long n, uint64_t seed)
{
    uint64_t st = seed;
    for (long i = 0; i < n; i++) {
        st += 0x9e3779b97f4a7c15ULL;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        v[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    }
}

int tn_eig_lowest(tn_ctx* c, int64_t N, int nev, double tol, int maxit, double* evals, double* evec,
                  int* n_matvec)
{
    if (!c) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(c->bound, TN_ERR_STATE, "tn_eig_lowest: no site bound");
    TN_REQUIRE(N == (int64_t)c->b_d * c->b_D1 * c->b_D2 && evals && evec && nev >= 1 && nev <= 2 && maxit >= 1,
               TN_ERR_ARG, "tn_eig_lowest: bad arguments");
    TN_REQUIRE(nev == 1 || N >= 2, TN_ERR_ARG, "tn_eig_lowest: a second eigenpair needs N >= 2");
    HeffOp op = bound_op(c);
    const long n = op.n();
    std::vector<double> start(evec, evec + N);
    bool any_nonzero = false, any_bad = false;
    for (long i = 0; i < N; i++) {
        if (!std::isfinite(start[i])) any_bad = true;
        else if (start[i] != 0.0) any_nonzero = true;
    }
    if (any_bad || !any_nonzero) random_start(start.data(), N, 0x5eedULL);
    psi_to_dev(c, start.data(), c->b_x.as<double>());
    LanczosResult r = lanczos(c, bound_cb, &op, n, c->b_x.as<double>(), tol, maxit, 4);
    int nmv = r.n_matvec;
    bool conv = r.converged;
    sync(c);
    psi_from_dev(c, c->b_x.as<double>(), evec);
    evals[0] = r.e0;
    if (nev > 1) {
        // second eigenpair (LapackEigenSolver.C:62-109 returns IL=1..IU=2): Lanczos
        // on the deflated operator from a start orthogonal to x0 up to rounding;
        // the shift is raised if the solve lands on the deflated vector
        const long pcap = mdot_part_size(n, 1);
        c->defl.ensure((size_t)(pcap + 8) * 8);
        DBuf x1;
        x1.ensure((size_t)n * 8);
        std::vector<double> s1(N);
        random_start(s1.data(), N, 0x5eed2ULL);
        double sigma = 2.0 * std::max(r.e1 - r.e0, 0.0) + 2.0 * std::fabs(r.e0) + 1.0;
        LanczosResult r2;
        for (int attempt = 0; attempt < 4; attempt++) {
            DeflatedOp dop{&op, c->b_x.as<double>(), sigma, c->defl.as<double>(), c->defl.as<double>() + 8, (int)pcap,
                           n};
            psi_to_dev(c, s1.data(), x1.as<double>());
            r2 = lanczos(c, deflated_cb, &dop, n, x1.as<double>(), tol, maxit, 4);
            nmv += r2.n_matvec;
            // overlap with x0: a large one means the shift did not lift x0 above lambda_2
            TN_HIP(mdot(c->b_x.as<double>(), n, 1, x1.as<double>(), n, c->defl.as<double>(), c->defl.as<double>() + 8,
                        (int)pcap, c->stream));
            double ov = 0;
            d2h(c, &ov, c->defl.p, 8);
            if (std::fabs(ov) < 1e-3) break;
            sigma *= 10.0;
        }
        // the true residual |H x1 - e1 x1| of the undeflated operator
        {
            DBuf hx;
            hx.ensure((size_t)n * 8);
            heff_apply(c, op, x1.as<double>(), hx.as<double>(), false);
            TN_HIP(axpby(-r2.e0, x1.as<double>(), 1.0, hx.as<double>(), n, c->stream));
            TN_HIP(mdot(hx.as<double>(), n, 1, hx.as<double>(), n, c->defl.as<double>(), c->defl.as<double>() + 8,
                        (int)pcap, c->stream));
            double r2n = 0;
            d2h(c, &r2n, c->defl.p, 8);
            hx.release();
            if (!(std::sqrt(std::max(r2n, 0.0)) <= 100.0 * std::max(tol, 1e-14 * std::fabs(sigma)) *
                                                     std::max(1.0, std::fabs(r2.e0))))
                r2.converged = false;
        }
        conv = conv && r2.converged;
        sync(c);
        psi_from_dev(c, x1.as<double>(), evec + N);
        x1.release();
        evals[1] = r2.e0;
    }
    if (n_matvec) *n_matvec = nmv;
    return conv ? TN_OK : TN_WARN_NOT_CONVERGED;
    ABI_END(c)
}

int tn_heff_apply_device(tn_ctx* c, const double* x, double* y, int64_t N)
{
    if (!c) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(c->bound, TN_ERR_STATE, "tn_heff_apply_device: no site bound");
    TN_REQUIRE(x && y && N == (int64_t)c->b_d * c->b_D1 * c->b_D2, TN_ERR_ARG,
               "tn_heff_apply_device: N != d*D1*D2");
    const int d = c->b_d, D1 = c->b_D1, D2 = c->b_D2, D1p = pad32(D1), D2p = pad32(D2);
    TN_HIP(psi_ref_to_int(x, c->b_x.as<double>(), d, D1, D2, D1p, D2p, c->stream));
    HeffOp op = bound_op(c);
    heff_apply(c, op, c->b_x.as<double>(), c->b_y.as<double>(), false);
    TN_HIP(psi_int_to_ref(c->b_y.as<double>(), y, d, D1, D2, D2p, c->stream));
    sync(c);
    return TN_OK;
    ABI_END(c)
}

int tn_heff_apply_device_noalloc(tn_ctx* c, const double* x, double* y, int64_t N)
{
    if (!c) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(c->bound, TN_ERR_STATE, "tn_heff_apply_device_noalloc: no site bound");
    TN_REQUIRE(x && y && N == (int64_t)c->b_d * c->b_D1 * c->b_D2, TN_ERR_ARG,
               "tn_heff_apply_device_noalloc: N != d*D1*D2");
    TN_REQUIRE(c->noalloc != nullptr, TN_ERR_STATE, "tn_heff_apply_device_noalloc: no work space on this device");
    const int d = c->b_d, D1 = c->b_D1, D2 = c->b_D2, D1p = pad32(D1), D2p = pad32(D2);
    HeffOp op = bound_op(c);
    // work space layout: x and y in the internal layout, then the stage buffers;
    // every byte a call reads is written earlier in the same call
    size_t b1 = 0, b2 = 0, bx = 0;
    heff_scratch_bytes(op, &b1, &b2, &bx);
    const size_t bv = ((size_t)op.n() * 8 + 255) / 256 * 256;
    const size_t total = 2 * bv + b1 + b2 + bx;
    TN_REQUIRE(total <= kNoAllocBytes, TN_ERR_ARG,
               "tn_heff_apply_device_noalloc: this site needs " + std::to_string(total) +
                   " bytes of work space, above the fixed " + std::to_string(kNoAllocBytes) + " bytes");
    char* base = (char*)c->noalloc;
    double* xi = (double*)base;
    double* yi = (double*)(base + bv);
    HeffScratch scr{base + 2 * bv, b1, base + 2 * bv + b1, b2, bx ? base + 2 * bv + b1 + b2 : nullptr, bx};
    TN_HIP(psi_ref_to_int(x, xi, d, D1, D2, D1p, D2p, c->stream));
    heff_apply(c, op, xi, yi, false, &scr);
    TN_HIP(psi_int_to_ref(yi, y, d, D1, D2, D2p, c->stream));
    sync(c);
    return TN_OK;
    ABI_END(c)
}

int tn_qr(tn_ctx* c, int mode, int m, int n, const double* A, double* Q, double* R)
{
    if (!c) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(A && Q && R && m > 0 && n > 0 && (mode == 0 || mode == 1), TN_ERR_ARG, "tn_qr: bad arguments");
    TN_REQUIRE(mode == 0 ? m >= n : m <= n, TN_ERR_ARG,
               mode == 0 ? "tn_qr: the QR form needs m >= n (use the LQ form)"
                         : "tn_qr: the LQ form needs m <= n (use the QR form)");
    // X (rows x cols, row-major, padded) = A (QR) or A^T (LQ); X = Q' Lacc^T
    const int rows = mode == 0 ? m : n, cols = mode == 0 ? n : m;
    const int rp = pad32(rows), cp = pad32(cols);
    TN_REQUIRE(rp >= cp, TN_ERR_ARG, "tn_qr: padded shapes");
    std::vector<double> h((size_t)rp * cp, 0.0);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            const double v = A[i + (size_t)j * m];
            if (mode == 0) h[(size_t)i * cp + j] = v;
            else h[(size_t)j * cp + i] = v;
        }
    DBuf a, q, l;
    a.ensure(h.size() * 8);
    q.ensure(h.size() * 8);
    l.ensure((size_t)cp * cp * 8);
    h2d(c, a.p, h.data(), h.size() * 8);
    int sh = 0;
    const bool ok = cholqr_device(c, a.as<double>(), cp, rp, cp, cols, q.as<double>(), l.as<double>(), 1e-11, &sh, true);
    d2h(c, h.data(), q.p, h.size() * 8);
    std::vector<double> hl((size_t)cp * cp);
    d2h(c, hl.data(), l.p, hl.size() * 8);
    a.release();
    q.release();
    l.release();
    if (mode == 0) {
        // Q (m x n) column-major, R (n x n) = Lacc^T upper, column-major
        for (int i = 0; i < m; i++)
            for (int j = 0; j < n; j++) Q[i + (size_t)j * m] = h[(size_t)i * cp + j];
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) R[i + (size_t)j * n] = hl[(size_t)j * cp + i];
    } else {
        // A = L Q: Q (m x n) = Q'^T, L (m x m) = Lacc lower, column-major
        for (int i = 0; i < m; i++)
            for (int j = 0; j < n; j++) Q[i + (size_t)j * m] = h[(size_t)j * cp + i];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) R[i + (size_t)j * m] = hl[(size_t)i * cp + j];
    }
    return ok ? TN_OK : TN_WARN_NOT_CONVERGED;
    ABI_END(c)
}

int tn_publish_test(tn_ctx* c, int n, const double* in, double* out)
{
    if (!c || !in || !out || n < 1 || n > TN_PUBLISH_MAX) return TN_ERR_ARG;
    ABI_BEGIN
    DBuf b;
    b.ensure((size_t)n * 8);
    h2d(c, b.p, in, (size_t)n * 8);
    d2h_small(c, out, b.as<double>(), n);
    b.release();
    return TN_OK;
    ABI_END(c)
}

int tn_svd(tn_ctx* c, int m, int n, const double* A, double* U, double* s, double* Vt)
{
    if (!c) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(m > 0 && n > 0 && A && U && s && Vt, TN_ERR_ARG, "tn_svd: bad arguments");
    const int k = std::min(m, n);
    c->svd_A.ensure((size_t)m * n * 8);
    h2d(c, c->svd_A.p, A, (size_t)m * n * 8);
    // A column-major: A(i,j) = A[i + j*m]
    tn::DBuf Ub, Vb;
    Ub.ensure((size_t)m * k * 8);
    Vb.ensure((size_t)k * n * 8);
    std::vector<double> sv;
    svd_device(c, c->svd_A.as<double>(), 1, m, m, n, Ub.as<double>(), k, m, Vb.as<double>(), n, k, sv);
    std::vector<double> hu((size_t)m * k), hv((size_t)k * n);
    d2h(c, hu.data(), Ub.p, hu.size() * 8);
    d2h(c, hv.data(), Vb.p, hv.size() * 8);
    Ub.release();
    Vb.release();
    for (int i = 0; i < m; i++)
        for (int j = 0; j < k; j++) U[i + (size_t)j * m] = hu[(size_t)i * k + j];
    for (int i = 0; i < k; i++)
        for (int j = 0; j < n; j++) Vt[i + (size_t)j * k] = hv[(size_t)i * n + j];
    for (int j = 0; j < k; j++) s[j] = sv[j];
    return TN_OK;
    ABI_END(c)
}

// ---- chain ------------------------------------------------------------------
int tn_chain_create(tn_ctx* c, int L, int d, const int* dw1, const int* dw2, const double* wcat, int D,
                    int Dcap, tn_chain** out)
{
    if (!c || !out) return TN_ERR_ARG;
    *out = nullptr;
    tn_chain* ch = new tn_chain();
    ABI_BEGIN
    TN_REQUIRE(L >= 1 && d >= 2 && D >= 1 && dw1 && dw2 && wcat, TN_ERR_ARG, "tn_chain_create: bad arguments");
    if (Dcap < D) Dcap = D;
    ch->ctx = c;
    ch->L = L;
    ch->d = d;
    ch->Dcap = Dcap;
    ch->Dcapp = pad32(Dcap);
    ch->dw1.assign(dw1, dw1 + L);
    ch->dw2.assign(dw2, dw2 + L);
    size_t off = 0;
    ch->W.resize(L);
    ch->Dwmax = 1;
    for (int i = 0; i < L; i++) {
        TN_REQUIRE(dw1[i] >= 1 && dw2[i] >= 1, TN_ERR_ARG, "bad MPO dims");
        if (i > 0) TN_REQUIRE(dw1[i] == dw2[i - 1], TN_ERR_ARG, "MPO bond dims do not match");
        size_t n = (size_t)dw1[i] * dw2[i] * d * d;
        ch->W[i].assign(wcat + off, wcat + off + n);
        off += n;
        ch->Dwmax = std::max(ch->Dwmax, std::max(dw1[i], dw2[i]));
    }
    TN_REQUIRE(dw1[0] == 1 && dw2[L - 1] == 1, TN_ERR_ARG, "edge MPO blocks must be 1 x Dw and Dw x 1");
    ch->site_cap = (size_t)ch->Dcapp * d * ch->Dcapp;
    ch->env_cap = (size_t)ch->Dcapp * ch->Dwmax * ch->Dcapp;
    ch->sites_buf.ensure(ch->site_cap * 8 * L);
    ch->spare_buf.ensure(ch->site_cap * 8);
    ch->env_buf.ensure(ch->env_cap * 8 * 2 * L);
    ch->bnd_buf.ensure((size_t)2 * 32 * 32 * ch->Dwmax * 8 + 512);
    ch->A.resize(L);
    ch->Lenv.resize(L);
    ch->Renv.resize(L);
    for (int i = 0; i < L; i++) {
        ch->A[i] = ch->sites_buf.as<double>() + (size_t)i * ch->site_cap;
        ch->Lenv[i] = ch->env_buf.as<double>() + (size_t)(2 * i) * ch->env_cap;
        ch->Renv[i] = ch->env_buf.as<double>() + (size_t)(2 * i + 1) * ch->env_cap;
    }
    ch->spare = ch->spare_buf.as<double>();
    TN_HIP(fill_zero(ch->sites_buf.p, (long)ch->sites_buf.bytes, c->stream));
    TN_HIP(fill_zero(ch->env_buf.p, (long)ch->env_buf.bytes, c->stream));
    TN_HIP(fill_zero(ch->bnd_buf.p, (long)ch->bnd_buf.bytes, c->stream));
    ch->bL = ch->bnd_buf.as<double>();
    ch->bR = ch->bL + 32 * 32 * ch->Dwmax + 32;
    const double one = 1.0;
    h2d(c, ch->bL, &one, 8);
    h2d(c, ch->bR, &one, 8);
    ch->mixH.resize(L);
    ch->mixL.resize(L);
    ch->mixR.resize(L);
    ch->mix2H.resize(L);
    ch->W12.resize(L);
    // MPSImp.C:69-90 shapes; Bond.C:11-25 lambda = 1/D
    ch->D1.resize(L);
    ch->D2.resize(L);
    for (int i = 0; i < L; i++) {
        ch->D1[i] = (i == 0) ? 1 : D;
        ch->D2[i] = (i == L - 1) ? 1 : D;
    }
    ch->lam.assign(std::max(L - 1, 1), std::vector<double>(D, 1.0 / D));
    ch->maxDelta.assign(std::max(L - 1, 1), 0.0);
    ch->compErr.assign(std::max(L - 1, 1), -99.0);
    ch->Emin.assign(L, 0.0);
    ch->IterDE.assign(L, 1.0);
    ch->GapE.assign(L, 0.0);
    sync(c);
    *out = ch;
    return TN_OK;
    }
    catch (const TnError& e) {
        int code = fail(c, e);
        delete ch;
        return code;
    }
}

int tn_chain_destroy(tn_chain* ch)
{
    if (!ch) return TN_OK;
    if (ch->ctx && ch->ctx->stream) (void)hipStreamSynchronize(ch->ctx->stream);
    ch->sites_buf.release();
    ch->spare_buf.release();
    ch->env_buf.release();
    drop_lo(ch);
    ch->bnd_buf.release();
    ch->theta.release();
    ch->w12buf.release();
    for (auto* v : {&ch->mixH, &ch->mixL, &ch->mixR, &ch->mix2H})
        for (auto& m : *v) m.buf.release();
    delete ch;
    return TN_OK;
}

int tn_chain_initialize(tn_chain* ch, int state, uint64_t seed)
{
    if (!ch) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    // MPSSite.C:84-140 per site, sign alternating
    uint64_t st = seed;
    int sgn = 1;
    const int d = ch->d, L = ch->L;
    for (int i = 0; i < L; i++) {
        const int D1 = ch->D1[i], D2 = ch->D2[i];
        std::vector<double> M((size_t)d * D1 * D2, 0.0);
        if (state == 1) {
            const double sc = 1.0 / std::sqrt((double)d * D1 * D2);
            for (int m = 0; m < d; m++)
                for (int i2 = 0; i2 < D2; i2++)
                    for (int i1 = 0; i1 < D1; i1++) {
                        st += 0x9e3779b97f4a7c15ULL;
                        uint64_t z = st;
                        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
                        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
                        z ^= z >> 31;
                        M[psi_ref(d, D1, m, i1, i2)] = ((double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * sc;
                    }
        } else if (state == 2) {
            if (sgn == 1) M[psi_ref(d, D1, 0, 0, 0)] = 1.0;
            else M[psi_ref(d, D1, d - 1, 0, 0)] = 1.0;
        } else if (state == 0) {
            if (i == 0 && L > 1) {
                for (int m = 0; m < d && m < D2; m++) M[psi_ref(d, D1, m, 0, m)] = sgn;
            } else if (i == L - 1 && L > 1) {
                for (int m = 0; m < d && m < D1; m++) M[psi_ref(d, D1, m, m, 0)] = sgn;
            } else {
                for (int m = 0; m < d; m++)
                    for (int k = 0; k < std::min(D1, D2); k++) M[psi_ref(d, D1, m, k, k)] = sgn / std::sqrt((double)d);
            }
        } else {
            TN_REQUIRE(false, TN_ERR_ARG, "unknown initial state");
        }
        upload_site(ch, i, D1, D2, M.data());
        sgn = -sgn;
    }
    return TN_OK;
    ABI_END(c)
}

int tn_chain_get_dims(const tn_chain* ch, int* D1, int* D2)
{
    if (!ch || !D1 || !D2) return TN_ERR_ARG;
    for (int i = 0; i < ch->L; i++) {
        D1[i] = ch->D1[i];
        D2[i] = ch->D2[i];
    }
    return TN_OK;
}

int tn_chain_get_site(tn_chain* ch, int site, double* M)
{
    if (!ch || !M || site < 0 || site >= ch->L) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    std::vector<double> v;
    download_site(ch, site, v);
    std::memcpy(M, v.data(), v.size() * 8);
    return TN_OK;
    ABI_END(c)
}

int tn_chain_set_site(tn_chain* ch, int site, int D1, int D2, const double* M)
{
    if (!ch || !M || site < 0 || site >= ch->L || D1 < 1 || D2 < 1) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    upload_site(ch, site, D1, D2, M);
    return TN_OK;
    ABI_END(c)
}

int tn_chain_normalize(tn_chain* ch, int dir)
{
    if (!ch || (dir != 0 && dir != 1)) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    if (dir == 0)
        for (int i = 0; i < ch->L; i++) svd_normalize(ch, i, 0, nullptr);
    else
        for (int i = ch->L - 1; i >= 0; i--) svd_normalize(ch, i, 1, nullptr);
    sync(c);
    return TN_OK;
    ABI_END(c)
}

int tn_chain_normalize_qr(tn_chain* ch, int dir)
{
    if (!ch || (dir != 0 && dir != 1)) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    if (dir == 0)
        for (int i = 0; i < ch->L; i++) svd_normalize(ch, i, 0, nullptr, 1);
    else
        for (int i = ch->L - 1; i >= 0; i--) svd_normalize(ch, i, 1, nullptr, 1);
    sync(c);
    return TN_OK;
    ABI_END(c)
}

int tn_chain_load_caches(tn_chain* ch, const tn_sweep_opts* opts)
{
    if (!ch) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    tn_sweep_opts o;
    tn_sweep_opts_default(&o);
    if (opts) o = *opts;
    for (int i = 0; i < ch->L - 1; i++) env_update(ch, i, 0, o, nullptr);
    for (int i = ch->L - 1; i >= 1; i--) env_update(ch, i, 1, o, nullptr);
    sync(c);
    return TN_OK;
    ABI_END(c)
}

int tn_chain_refine_site(tn_chain* ch, int site, int dir, const tn_sweep_opts* opts, tn_sweep_stats* st)
{
    if (!ch || site < 0 || site >= ch->L || (dir != 0 && dir != 1)) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    tn_sweep_opts o;
    tn_sweep_opts_default(&o);
    if (opts) o = *opts;
    refine_site(ch, site, dir, o, st);
    return TN_OK;
    ABI_END(c)
}

int tn_chain_sweep(tn_chain* ch, int dir, const tn_sweep_opts* opts, tn_sweep_stats* st)
{
    if (!ch || (dir != 0 && dir != 1)) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    tn_sweep_opts o;
    tn_sweep_opts_default(&o);
    if (opts) o = *opts;
    double t0 = now_ms();
    if (dir == 0)
        for (int i = 0; i < ch->L; i++) refine_site(ch, i, 0, o, st);
    else
        for (int i = ch->L - 1; i >= 0; i--) refine_site(ch, i, 1, o, st);
    sync(c);
    if (st) st->wall_ms += now_ms() - t0;
    return TN_OK;
    ABI_END(c)
}

int tn_chain_increase_bond_dims(tn_chain* ch, int D)
{
    if (!ch || D < 1) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    chain_reserve(ch, D);
    // MPSImp.C:123-147
    const int L = ch->L, d = ch->d;
    long Dideal = 1;
    int ia = 0;
    auto grow_bond = [&](int b, int nD) {  // Bond.C:60-84
        std::vector<double>& l = ch->lam[b];
        const int oD = (int)l.size();
        std::vector<double> nl(nD);
        if (nD > oD) {
            double mn = l.empty() ? 0.0 : *std::min_element(l.begin(), l.end());
            for (int k = 0; k < oD; k++) nl[k] = l[k];
            for (int k = oD; k < nD; k++) nl[k] = mn / 10.0;
        } else {
            for (int k = 0; k < nD; k++) nl[k] = l[k];
        }
        double nr = 0;
        for (double v : nl) nr += v * v;
        nr = std::sqrt(nr);
        for (double& v : nl) v = nr > 0 ? v / nr : v;
        l = nl;
    };
    for (; ia < L / 2; ia++) {
        int D1 = (int)std::min<long>(Dideal, D), D2 = (int)std::min<long>(Dideal * d, D);
        site_resize(ch, ia, D1, D2, true);
        grow_bond(ia, D2);
        Dideal *= d;
    }
    Dideal = 1;
    for (int ib = L - 1; ib >= ia; ib--) {
        int D2 = (int)std::min<long>(Dideal, D), D1 = (int)std::min<long>(Dideal * d, D);
        site_resize(ch, ib, D1, D2, true);
        if (ib >= 1) grow_bond(ib - 1, D1);
        Dideal *= d;
    }
    return TN_OK;
    ABI_END(c)
}

int tn_chain_deltas(const tn_chain* ch, double* max_de, double* max_dl)
{
    if (!ch) return TN_ERR_ARG;
    double de = 0, dl = 0;
    for (int i = 0; i < ch->L; i++) de = std::max(de, std::fabs(ch->IterDE[i]));
    for (int b = 0; b < ch->L - 1; b++) dl = std::max(dl, ch->maxDelta[b]);
    if (max_de) *max_de = de;
    if (max_dl) *max_dl = dl;
    return TN_OK;
}

int tn_chain_bond_sv(const tn_chain* ch, int b, double* sv, int cap)
{
    if (!ch || b < 0 || b >= ch->L - 1) return TN_ERR_ARG;
    const std::vector<double>& l = ch->lam[b];
    int n = (int)l.size();
    for (int k = 0; k < n && k < cap; k++) sv[k] = l[k];
    return n;
}

int tn_chain_norm_status(tn_chain* ch, double eps, char* out)
{
    if (!ch || !out || !(eps > 0.0)) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    // MPSSite.C:173-211: 'A' sum_{m,a} A*A = 1, 'B' sum_{m,b} A A* = 1, 'I' both, 'M' neither
    const int d = ch->d;
    DBuf g, dev, part;
    g.ensure((size_t)ch->Dcapp * ch->Dcapp * 8);
    dev.ensure(2 * 8);
    part.ensure(1024 * 8);
    for (int i = 0; i < ch->L; i++) {
        const int D1 = ch->D1[i], D2 = ch->D2[i], D1p = pad32(D1), D2p = pad32(D2);
        GemmDesc gl{};
        gl.tin = T_F64; gl.tout = T_F64; gl.a_kmajor = false; gl.b_kmajor = false;
        gl.M = D2p; gl.N = D2p; gl.K = D1p * d; gl.batch = 1;
        gl.A = ch->A[i]; gl.lda = D2p; gl.B = ch->A[i]; gl.ldb = D2p; gl.C = g.p; gl.ldc = D2p; gl.alpha = 1.0f;
        TN_HIP(gemm(gl, c->stream));
        TN_HIP(eye_deviation(g.as<double>(), D2p, D2, dev.as<double>(), part.as<double>(), c->stream));
        GemmDesc gr{};
        gr.tin = T_F64; gr.tout = T_F64; gr.a_kmajor = true; gr.b_kmajor = true;
        gr.M = D1p; gr.N = D1p; gr.K = d * D2p; gr.batch = 1;
        gr.A = ch->A[i]; gr.lda = (long)d * D2p; gr.B = ch->A[i]; gr.ldb = (long)d * D2p; gr.C = g.p; gr.ldc = D1p;
        gr.alpha = 1.0f;
        TN_HIP(gemm(gr, c->stream));
        TN_HIP(eye_deviation(g.as<double>(), D1p, D1, dev.as<double>() + 1, part.as<double>(), c->stream));
        double e[2];
        d2h(c, e, dev.p, 16);
        const bool A = e[0] < eps, B = e[1] < eps;
        out[i] = (A && B) ? 'I' : (A ? 'A' : (B ? 'B' : 'M'));
    }
    out[ch->L] = 0;
    g.release();
    dev.release();
    part.release();
    return TN_OK;
    ABI_END(c)
}

// ---- checkpoint / resume ------------------------------------------------------
// File layout (little-endian): "TNCHKPT1", int32 version, L, d, Dcap, flags,
// dw1[L], dw2[L], float64 W (site order); per site int32 D1, D2 and the site
// tensor in the reference layout; per bond int32 n, lambda[n], maxDelta,
// compErr; per site Emin, IterDE, GapE; with flags & 1 the FP64 environments
// (left env after site i: D2 x Dw2 x D2 as [bra][w][ket]; right env from site i:
// Dw1 x D1 x D1 as [w][bra][ket]); int64 meta length and the caller's meta
// bytes; uint64 FNV-1a hash of everything before it.
namespace {
struct Fnv {
    uint64_t h = 1469598103934665603ULL;
    void add(const void* p, size_t n)
    {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; i++) {
            h ^= b[i];
            h *= 1099511628211ULL;
        }
    }
};
struct Writer {
    FILE* f;
    Fnv h;
    void put(const void* p, size_t n)
    {
        TN_REQUIRE(fwrite(p, 1, n, f) == n, TN_ERR_STATE, "tn_chain_save: write failed");
        h.add(p, n);
    }
    void i32(int v) { put(&v, 4); }
    void f64(const double* p, size_t n) { put(p, n * 8); }
};
struct Reader {
    FILE* f;
    Fnv h;
    void get(void* p, size_t n)
    {
        TN_REQUIRE(fread(p, 1, n, f) == n, TN_ERR_ARG, "tn_chain_load: truncated file");
        h.add(p, n);
    }
    int i32()
    {
        int v;
        get(&v, 4);
        return v;
    }
    void f64(double* p, size_t n) { get(p, n * 8); }
};
struct FileCloser {
    FILE* f;
    ~FileCloser()
    {
        if (f) fclose(f);
    }
};
const char kCkptMagic[8] = {'T', 'N', 'C', 'H', 'K', 'P', 'T', '1'};
}  // namespace

int tn_chain_save(tn_chain* ch, const char* path, const void* meta, int64_t meta_bytes, int flags)
{
    if (!ch || !path || meta_bytes < 0 || (meta_bytes > 0 && !meta) || (flags & ~1)) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    const std::string tmp = std::string(path) + ".tmp";
    FileCloser fc{fopen(tmp.c_str(), "wb")};
    TN_REQUIRE(fc.f, TN_ERR_ARG, std::string("tn_chain_save: cannot open ") + tmp);
    Writer w{fc.f};
    const int L = ch->L, d = ch->d;
    w.put(kCkptMagic, 8);
    w.i32(1);
    w.i32(L);
    w.i32(d);
    w.i32(ch->Dcap);
    w.i32(flags);
    for (int i = 0; i < L; i++) w.i32(ch->dw1[i]);
    for (int i = 0; i < L; i++) w.i32(ch->dw2[i]);
    for (int i = 0; i < L; i++) w.f64(ch->W[i].data(), ch->W[i].size());
    std::vector<double> M;
    for (int i = 0; i < L; i++) {
        w.i32(ch->D1[i]);
        w.i32(ch->D2[i]);
        download_site(ch, i, M);
        w.f64(M.data(), M.size());
    }
    for (int b = 0; b < L - 1; b++) {
        w.i32((int)ch->lam[b].size());
        w.f64(ch->lam[b].data(), ch->lam[b].size());
        w.f64(&ch->maxDelta[b], 1);
        w.f64(&ch->compErr[b], 1);
    }
    for (int i = 0; i < L; i++) {
        w.f64(&ch->Emin[i], 1);
        w.f64(&ch->IterDE[i], 1);
        w.f64(&ch->GapE[i], 1);
    }
    if (flags & 1) {
        for (int i = 0; i < L; i++) {
            // logical blocks only: left [D2][Dw2][D2], right [Dw1][D1][D1]
            const int D1 = ch->D1[i], D2 = ch->D2[i], D1p = pad32(D1), D2p = pad32(D2);
            const int Dw1 = ch->dw1[i], Dw2 = ch->dw2[i];
            std::vector<double> buf((size_t)std::max(D2p * Dw2 * D2p, Dw1 * D1p * D1p));
            std::vector<double> out;
            d2h(c, buf.data(), ch->Lenv[i], (size_t)D2p * Dw2 * D2p * 8);
            out.clear();
            for (int a = 0; a < D2; a++)
                for (int x = 0; x < Dw2; x++)
                    for (int b = 0; b < D2; b++) out.push_back(buf[((size_t)a * Dw2 + x) * D2p + b]);
            w.f64(out.data(), out.size());
            d2h(c, buf.data(), ch->Renv[i], (size_t)Dw1 * D1p * D1p * 8);
            out.clear();
            for (int x = 0; x < Dw1; x++)
                for (int a = 0; a < D1; a++)
                    for (int b = 0; b < D1; b++) out.push_back(buf[((size_t)x * D1p + a) * D1p + b]);
            w.f64(out.data(), out.size());
        }
    }
    const int64_t mb = meta_bytes;
    w.put(&mb, 8);
    if (mb > 0) w.put(meta, (size_t)mb);
    const uint64_t hash = w.h.h;
    TN_REQUIRE(fwrite(&hash, 1, 8, fc.f) == 8, TN_ERR_STATE, "tn_chain_save: write failed");
    TN_REQUIRE(fclose(fc.f) == 0, TN_ERR_STATE, "tn_chain_save: close failed");
    fc.f = nullptr;
    TN_REQUIRE(std::rename(tmp.c_str(), path) == 0, TN_ERR_STATE, std::string("tn_chain_save: cannot rename to ") + path);
    return TN_OK;
    ABI_END(c)
}

int tn_chain_load(tn_ctx* c, const char* path, int Dcap, tn_chain** out, void* meta, int64_t meta_cap,
                  int64_t* meta_bytes)
{
    if (!c || !path || !out || meta_cap < 0 || (meta_cap > 0 && !meta)) return TN_ERR_ARG;
    *out = nullptr;
    tn_chain* ch = nullptr;
    ABI_BEGIN
    FileCloser fc{fopen(path, "rb")};
    TN_REQUIRE(fc.f, TN_ERR_ARG, std::string("tn_chain_load: cannot open ") + path);
    Reader r{fc.f};
    char magic[8];
    r.get(magic, 8);
    TN_REQUIRE(std::memcmp(magic, kCkptMagic, 8) == 0, TN_ERR_ARG, "tn_chain_load: not a chain checkpoint");
    TN_REQUIRE(r.i32() == 1, TN_ERR_ARG, "tn_chain_load: unknown version");
    const int L = r.i32(), d = r.i32(), fDcap = r.i32(), flags = r.i32();
    TN_REQUIRE(L >= 1 && L <= (1 << 20) && d >= 2 && d <= 64 && fDcap >= 1 && (flags & ~1) == 0, TN_ERR_ARG,
               "tn_chain_load: bad header");
    std::vector<int> dw1(L), dw2(L);
    for (int i = 0; i < L; i++) dw1[i] = r.i32();
    for (int i = 0; i < L; i++) dw2[i] = r.i32();
    size_t nw = 0;
    for (int i = 0; i < L; i++) {
        TN_REQUIRE(dw1[i] >= 1 && dw2[i] >= 1 && dw1[i] <= 4096 && dw2[i] <= 4096, TN_ERR_ARG,
                   "tn_chain_load: bad MPO dims");
        nw += (size_t)dw1[i] * dw2[i] * d * d;
    }
    std::vector<double> wcat(nw);
    r.f64(wcat.data(), nw);
    std::vector<int> D1(L), D2(L);
    std::vector<std::vector<double>> M(L);
    int Dmax = 1;
    for (int i = 0; i < L; i++) {
        D1[i] = r.i32();
        D2[i] = r.i32();
        TN_REQUIRE(D1[i] >= 1 && D2[i] >= 1 && D1[i] <= fDcap && D2[i] <= fDcap, TN_ERR_ARG,
                   "tn_chain_load: bad site dims");
        Dmax = std::max(Dmax, std::max(D1[i], D2[i]));
        M[i].resize((size_t)d * D1[i] * D2[i]);
        r.f64(M[i].data(), M[i].size());
    }
    const int cap = std::max(std::max(Dcap, fDcap), Dmax);
    int rc = tn_chain_create(c, L, d, dw1.data(), dw2.data(), wcat.data(), Dmax, cap, &ch);
    if (rc != TN_OK) return rc;
    for (int i = 0; i < L; i++) upload_site(ch, i, D1[i], D2[i], M[i].data());
    for (int b = 0; b < L - 1; b++) {
        const int n = r.i32();
        TN_REQUIRE(n >= 0 && n <= fDcap, TN_ERR_ARG, "tn_chain_load: bad bond size");
        ch->lam[b].resize(n);
        r.f64(ch->lam[b].data(), n);
        r.f64(&ch->maxDelta[b], 1);
        r.f64(&ch->compErr[b], 1);
    }
    for (int i = 0; i < L; i++) {
        r.f64(&ch->Emin[i], 1);
        r.f64(&ch->IterDE[i], 1);
        r.f64(&ch->GapE[i], 1);
    }
    if (flags & 1) {
        for (int i = 0; i < L; i++) {
            const int a1 = D1[i], a2 = D2[i], D1p = pad32(a1), D2p = pad32(a2), w1 = dw1[i], w2 = dw2[i];
            std::vector<double> in((size_t)std::max(a2 * w2 * a2, w1 * a1 * a1));
            std::vector<double> buf((size_t)D2p * w2 * D2p, 0.0);
            r.f64(in.data(), (size_t)a2 * w2 * a2);
            for (int a = 0; a < a2; a++)
                for (int x = 0; x < w2; x++)
                    for (int b = 0; b < a2; b++) buf[((size_t)a * w2 + x) * D2p + b] = in[((size_t)a * w2 + x) * a2 + b];
            h2d(c, ch->Lenv[i], buf.data(), buf.size() * 8);
            buf.assign((size_t)w1 * D1p * D1p, 0.0);
            r.f64(in.data(), (size_t)w1 * a1 * a1);
            for (int x = 0; x < w1; x++)
                for (int a = 0; a < a1; a++)
                    for (int b = 0; b < a1; b++) buf[((size_t)x * D1p + a) * D1p + b] = in[((size_t)x * a1 + a) * a1 + b];
            h2d(c, ch->Renv[i], buf.data(), buf.size() * 8);
        }
    }
    int64_t mb = 0;
    r.get(&mb, 8);
    TN_REQUIRE(mb >= 0 && mb <= ((int64_t)1 << 30), TN_ERR_ARG, "tn_chain_load: bad meta length");
    std::vector<char> mbuf((size_t)mb);
    if (mb > 0) r.get(mbuf.data(), (size_t)mb);
    const uint64_t expect = r.h.h;
    uint64_t hash = 0;
    TN_REQUIRE(fread(&hash, 1, 8, fc.f) == 8 && hash == expect, TN_ERR_ARG, "tn_chain_load: checksum mismatch");
    if (mb > 0 && meta_cap > 0) std::memcpy(meta, mbuf.data(), (size_t)std::min(mb, meta_cap));
    if (meta_bytes) *meta_bytes = mb;
    *out = ch;
    return TN_OK;
    }
    catch (const TnError& e) {
        int code = fail(c, e);
        if (ch) tn_chain_destroy(ch);
        return code;
    }
}

int tn_chain_expectation(tn_chain* ch, const int* dw1, const int* dw2, const double* wcat, double* value)
{
    if (!ch || !dw1 || !dw2 || !wcat || !value) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    const int L = ch->L, d = ch->d;
    TN_REQUIRE(dw1[0] == 1 && dw2[L - 1] == 1, TN_ERR_ARG, "edge MPO blocks must be 1 x Dw and Dw x 1");
    int Dwm = 1;
    for (int i = 0; i < L; i++) Dwm = std::max(Dwm, std::max(dw1[i], dw2[i]));
    const size_t cap = (size_t)ch->Dcapp * Dwm * ch->Dcapp;
    tn::DBuf F0, F1;
    F0.ensure(std::max(cap, (size_t)32 * 32) * 8);
    F1.ensure(std::max(cap, (size_t)32 * 32) * 8);
    TN_HIP(fill_zero(F0.p, (long)F0.bytes, c->stream));
    const double one = 1.0;
    h2d(c, F0.p, &one, 8);
    double* Fin = F0.as<double>();
    double* Fout = F1.as<double>();
    size_t off = 0;
    DevMix mix;
    for (int i = 0; i < L; i++) {
        std::vector<double> Wi(wcat + off, wcat + off + (size_t)dw1[i] * dw2[i] * d * d);
        off += Wi.size();
        const int D1p = pad32(ch->D1[i]), D2p = pad32(ch->D2[i]);
        mix_build(mix, MIX_ENVL, Wi.data(), dw1[i], dw2[i], d, D1p, D2p, c->stream);
        env_left(c, TN_FP64, d, D1p, D2p, dw1[i], dw2[i], mix, Fin, nullptr, ch->A[i], Fout, 0);
        sync(c);
        mix.wsrc = nullptr;  // W buffer is local: never reuse by address
        std::swap(Fin, Fout);
    }
    d2h(c, value, Fin, 8);
    mix.buf.release();
    F0.release();
    F1.release();
    return TN_OK;
    ABI_END(c)
}

int tn_chain_sweep2(tn_chain* ch, int dir, int Dmax, double eps_sv, const tn_sweep_opts* opts,
                    tn_sweep_stats* st, double* max_trunc)
{
    if (!ch || (dir != 0 && dir != 1) || ch->L < 2) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    tn_sweep_opts o;
    tn_sweep_opts_default(&o);
    if (opts) o = *opts;
    TN_REQUIRE(Dmax >= 1, TN_ERR_ARG, "Dmax >= 1");
    if (Dmax > ch->Dcap) {
        chain_reserve(ch, Dmax);
        TN_REQUIRE(false, TN_ERR_STATE, "tn_chain_sweep2: Dmax grew the chain; reload caches first");
    }
    hipStream_t s = c->stream;
    const int d = ch->d, L = ch->L, p2 = d * d;
    double mtr = 0.0;
    double t0w = now_ms();
    for (int step = 0; step < L - 1; step++) {
        const int i = (dir == 0) ? step : L - 2 - step, j = i + 1;
        const int D1 = ch->D1[i], Dm = ch->D2[i], D3 = ch->D2[j];
        const int D1p = pad32(D1), Dmp = pad32(Dm), D3p = pad32(D3);
        const int Dw1 = ch->dw1[i], Dwm = ch->dw2[i], Dw3 = ch->dw2[j];
        // Theta[a][m1][m2][c] = A_i[(a,m1)][b] A_j[b][(m2,c)]
        const long N = (long)D1p * p2 * D3p;
        ch->theta.ensure((size_t)N * 8);
        GemmDesc g{};
        g.tin = T_F64; g.tout = T_F64; g.a_kmajor = true; g.b_kmajor = false;
        g.M = D1p * d; g.N = d * D3p; g.K = Dmp; g.batch = 1;
        g.A = ch->A[i]; g.lda = Dmp; g.B = ch->A[j]; g.ldb = (long)d * D3p;
        g.C = ch->theta.p; g.ldc = (long)d * D3p; g.alpha = 1.0f;
        gemm_run(c, g);
        two_site_W(d, Dw1, Dwm, Dw3, ch->W[i].data(), ch->W[j].data(), ch->W12[i]);
        HeffCall hc;
        hc.op.p = p2;
        hc.op.D1p = D1p;
        hc.op.D2p = D3p;
        hc.op.Dw1 = Dw1;
        hc.op.Dw2 = Dw3;
        hc.op.L = i > 0 ? ch->Lenv[i - 1] : ch->bL;
        hc.op.R = j < L - 1 ? ch->Renv[j + 1] : ch->bR;
        hc.op.precision = o.precision;
        hc.op.Llo = lo_left(ch, o.precision, i);
        hc.op.Rlo = lo_right(ch, o.precision, j);
        ch->mix2H[i].wsrc = nullptr;  // W12 rebuilt each visit
        mix_build(ch->mix2H[i], MIX_HEFF_T, ch->W12[i].data(), Dw1, Dw3, p2, D1p, D3p, s);
        hc.op.mix = &ch->mix2H[i];
        hc.op.tile = o.gemm_tile;
        hc.st = st;
        hc.timed = o.time_heff != 0;
        hc.D1 = D1;
        hc.D2 = D3;
        double tl = now_ms();
        LanczosResult r = lanczos(c, heff_cb, &hc, N, ch->theta.as<double>(), o.eig_tol, o.max_krylov, o.max_restarts,
                                  0.0, o.reorth == 1, unit_roundoff(o.precision), true);
        if (st) st->n_full_reorth += r.n_full;
        if (o.refine && o.refine_precision < o.precision) {
            HeffCall h64 = hc;
            h64.op.precision = o.refine_precision;
            h64.op.Llo = lo_left(ch, o.refine_precision, i);
            h64.op.Rlo = lo_right(ch, o.refine_precision, j);
            LanczosResult r2 = lanczos(c, heff_cb, &h64, N, ch->theta.as<double>(), o.refine_tol, o.max_krylov,
                                       o.max_restarts, 0.0, o.reorth == 1, unit_roundoff(o.refine_precision), true);
            if (st) st->n_full_reorth += r2.n_full;
            r.e0 = r2.e0;
            r.e1 = r2.e1;
            r.converged = r2.converged;
        }
        heff_timing_flush(c);
        if (st) {
            (r.converged ? st->n_solve_converged : st->n_solve_capped)++;
            st->lanczos_ms += now_ms() - tl;
            st->energy = r.e0;
        }
        ch->IterDE[i] = r.e0 - ch->Emin[i];
        ch->Emin[i] = r.e0;
        // SVD of Theta as (D1*d) x (d*D3): pack logical matrix
        double ts = now_ms();
        const int rr = D1 * d, cc = d * D3, k = std::min(rr, cc);
        c->svd_A.ensure((size_t)rr * cc * 8 + 64);
        TN_HIP(copy3(ch->theta.as<double>(), (long)d * D3p, D3p, 1, c->svd_A.as<double>(), rr, d, D3, D3, s));
        const int kp = pad32(k);
        c->svd_U.ensure((size_t)rr * kp * 8 + 64);
        c->svd_Vt.ensure((size_t)kp * cc * 8 + 64);
        std::vector<double> sv;
        svd_device(c, c->svd_A.as<double>(), cc, 1, rr, cc, c->svd_U.as<double>(), kp, rr,
                   c->svd_Vt.as<double>(), cc, kp, sv);
        double disc = 0, resc = 1;
        int Dk = tn_svcompress_rank(k, sv.data(), Dmax, eps_sv, &disc, &resc);
        if (Dk < 1) Dk = 1;
        std::vector<double> sk(sv.begin(), sv.begin() + Dk);
        if (Dk < k)
            for (double& v : sk) v *= resc;
        mtr = std::max(mtr, disc);
        bond_set_sv(ch, i, sk, disc);
        double nr = 0;
        for (double v : sk) nr += v * v;
        nr = std::sqrt(nr);
        const int Dkp = pad32(Dk);
        std::vector<double> scl(std::max(Dkp, 32), 0.0);
        for (int t = 0; t < Dk; t++) scl[t] = sk[t] / nr;
        c->pack.ensure((size_t)scl.size() * 8 + 64);
        h2d(c, c->pack.p, scl.data(), scl.size() * 8);
        if (dir == 0) {
            // A_i <- U[:, :Dk] ; A_j <- diag(s) Vt[:Dk]
            TN_HIP(copy2(c->svd_U.as<double>(), kp, 1, rr, Dk, ch->A[i], Dkp, 1, D1p * d, Dkp, s));
            TN_HIP(scale_rows(c->svd_Vt.as<double>(), cc, Dk, cc, c->pack.as<double>(), s));
            TN_HIP(copy3(c->svd_Vt.as<double>(), cc, D3, 1, ch->A[j], Dk, d, D3, D3p, s));
            if (Dkp > Dk) TN_HIP(fill_zero(ch->A[j] + (size_t)Dk * d * D3p, (long)(Dkp - Dk) * d * D3p * 8, s));
        } else {
            // A_j <- Vt[:Dk] ; A_i <- U diag(s)
            TN_HIP(copy3(c->svd_Vt.as<double>(), cc, D3, 1, ch->A[j], Dk, d, D3, D3p, s));
            if (Dkp > Dk) TN_HIP(fill_zero(ch->A[j] + (size_t)Dk * d * D3p, (long)(Dkp - Dk) * d * D3p * 8, s));
            TN_HIP(scale_cols(c->svd_U.as<double>(), kp, rr, Dk, c->pack.as<double>(), s));
            TN_HIP(copy2(c->svd_U.as<double>(), kp, 1, rr, Dk, ch->A[i], Dkp, 1, D1p * d, Dkp, s));
        }
        ch->D2[i] = Dk;
        ch->D1[j] = Dk;
        if (st) {
            sync(c);
            st->svd_ms += now_ms() - ts;
            st->n_svd++;
        }
        if (dir == 0) env_update(ch, i, 0, o, st);
        else env_update(ch, j, 1, o, st);
    }
    sync(c);
    if (st) st->wall_ms += now_ms() - t0w;
    if (max_trunc) *max_trunc = mtr;
    return TN_OK;
    ABI_END(c)
}

int tn_chain_heff_bench(tn_chain* ch, int site, int precision, int reps, int skip_identity, int env_precision,
                        double* s1_ms, double* mix_ms, double* s3_ms, double* s1_flops, double* s3_flops)
{
    if (!ch || site < 0 || site >= ch->L || reps < 1) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    ABI_BEGIN
    heff_timing_flush(c);  // the events are reused below
    int wl = -1, wr = -1;
    if (skip_identity) env_identity(ch, site, ident_tol(env_precision), &wl, &wr);
    HeffOp op = make_heff(ch, site, precision, 0, wl, wr);
    const long N = op.n();
    c->b_y.ensure((size_t)N * 8);
    double a = 0, b = 0, e = 0;
    heff_apply(c, op, ch->A[site], c->b_y.as<double>(), true);  // warm
    sync(c);
    for (int r = 0; r < reps; r++) {
        heff_apply(c, op, ch->A[site], c->b_y.as<double>(), true);
        double x1, x2, x3;
        heff_stage_times(c, &x1, &x2, &x3);
        a += x1;
        b += x2;
        e += x3;
    }
    const double D1 = ch->D1[site], D2 = ch->D2[site];
    if (s1_ms) *s1_ms = a / reps;
    if (mix_ms) *mix_ms = b / reps;
    if (s3_ms) *s3_ms = e / reps;
    if (s1_flops) *s1_flops = 2.0 * D1 * op.p * (op.Dw2 - (wr >= 0 ? 1 : 0)) * D2 * D2;
    if (s3_flops) *s3_flops = 2.0 * D1 * op.p * D2 * (op.Dw1 - (wl >= 0 ? 1 : 0)) * D1;
    return TN_OK;
    ABI_END(c)
}

// ---- diagnostics --------------------------------------------------------------
int tn_gemm_test(tn_ctx* c, int tin, int tout, int akm, int bkm, int M, int N, int K, int batch,
                 const double* A, const double* B, double* C, int tile, int ksplit)
{
    if (!c || !A || !B || !C) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(M > 0 && N > 0 && K > 0 && batch > 0, TN_ERR_ARG, "tn_gemm_test: bad dims");
    const long nA = (long)batch * M * K, nB = (long)batch * K * N, nC = (long)batch * M * N;
    tn::DBuf a64, b64, alo, blo, cout, c64;
    a64.ensure(nA * 8);
    b64.ensure(nB * 8);
    h2d(c, a64.p, A, nA * 8);
    h2d(c, b64.p, B, nB * 8);
    const void* Ad = a64.p;
    const void* Bd = b64.p;
    if (tin == T_BF16X3) {
        TN_REQUIRE(akm && bkm, TN_ERR_ARG, "tn_gemm_test: the split-bfloat16 GEMM takes K-major operands");
        alo.ensure(nA * 6);
        blo.ensure(nB * 6);
        TN_HIP(split3(T_F64, a64.p, alo.p, nA, nA, c->stream));
        TN_HIP(split3(T_F64, b64.p, blo.p, nB, nB, c->stream));
        Ad = alo.p;
        Bd = blo.p;
    } else if (tin != T_F64) {
        alo.ensure(nA * 4);
        blo.ensure(nB * 4);
        TN_HIP(convert(T_F64, a64.p, tin, alo.p, nA, c->stream));
        TN_HIP(convert(T_F64, b64.p, tin, blo.p, nB, c->stream));
        Ad = alo.p;
        Bd = blo.p;
    }
    cout.ensure(nC * 8);
    c64.ensure(nC * 8);
    GemmDesc g{};
    g.tin = tin; g.tout = tout; g.a_kmajor = akm != 0; g.b_kmajor = bkm != 0;
    g.M = M; g.N = N; g.K = K; g.batch = batch;
    g.A = Ad; g.lda = akm ? K : M; g.sA = (long)M * K;
    g.B = Bd; g.ldb = bkm ? K : N; g.sB = (long)K * N;
    g.C = cout.p; g.ldc = N; g.sC = (long)M * N;
    g.alpha = 1.0f; g.tile = tile;
    g.psA = nA; g.psB = nB;
    gemm_split(c, g, ksplit);
    TN_HIP(convert(tout, cout.p, T_F64, c64.p, nC, c->stream));
    d2h(c, C, c64.p, nC * 8);
    for (auto* b : {&a64, &b64, &alo, &blo, &cout, &c64}) b->release();
    return TN_OK;
    ABI_END(c)
}

int tn_copy_bench(tn_ctx* c, int64_t n, double* ms)
{
    if (!c || n <= 0) return TN_ERR_ARG;
    ABI_BEGIN
    tn::DBuf a, b;
    a.ensure((size_t)n * 8);
    b.ensure((size_t)n * 8);
    TN_HIP(fill_value(a.as<double>(), 1.0, n, c->stream));
    TN_HIP(scale_copy(1.0, a.as<double>(), b.as<double>(), n, c->stream));  // warm
    TN_HIP(hipEventRecord(c->ev[6], c->stream));
    TN_HIP(scale_copy(1.0, a.as<double>(), b.as<double>(), n, c->stream));
    TN_HIP(hipEventRecord(c->ev[7], c->stream));
    TN_HIP(hipEventSynchronize(c->ev[7]));
    float t = 0;
    TN_HIP(hipEventElapsedTime(&t, c->ev[6], c->ev[7]));
    if (ms) *ms = t;
    a.release();
    b.release();
    return TN_OK;
    ABI_END(c)
}

int tn_cholqr_test(tn_ctx* c, int m, int n, const double* A, double tol, double* Q, double* R, int* shifted,
                   int triangular)
{
    if (!c || !A || !Q || !R) return TN_ERR_ARG;
    ABI_BEGIN
    TN_REQUIRE(m > 0 && n > 0 && m >= n, TN_ERR_ARG, "tn_cholqr_test: need m >= n > 0");
    const int mp = pad32(m), np_ = pad32(n);
    TN_REQUIRE(mp >= np_, TN_ERR_ARG, "tn_cholqr_test: padded m < padded n");
    tn::DBuf a, q, l;
    a.ensure((size_t)mp * np_ * 8);
    q.ensure((size_t)mp * np_ * 8);
    l.ensure((size_t)np_ * np_ * 8);
    TN_HIP(fill_zero(a.p, (long)mp * np_ * 8, c->stream));
    std::vector<double> h((size_t)mp * np_, 0.0);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) h[(size_t)i * np_ + j] = A[(size_t)i * n + j];
    h2d(c, a.p, h.data(), h.size() * 8);
    int sh = 0;
    const bool ok =
        cholqr_device(c, a.as<double>(), np_, mp, np_, n, q.as<double>(), l.as<double>(), tol, &sh, triangular != 0);
    d2h(c, h.data(), q.p, h.size() * 8);
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) Q[(size_t)i * n + j] = h[(size_t)i * np_ + j];
    std::vector<double> hl((size_t)np_ * np_);
    d2h(c, hl.data(), l.p, hl.size() * 8);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) R[(size_t)i * n + j] = hl[(size_t)j * np_ + i];  // R = L^T
    if (shifted) *shifted = sh;
    for (auto* b : {&a, &q, &l}) b->release();
    return ok ? TN_OK : 1;
    ABI_END(c)
}

}  // extern "C"

