// tn_precond.hip -- Davidson local eigensolver with a learned diagonal
// preconditioner (no reference counterpart; the reference solves the local
// problem with dense LAPACK or ARPACK/PRIMME Lanczos).
//
// Correction of the Davidson step: t_i = g_i * s0_i * r_i, s0_i = -1/den_i,
// den_i = diag_i - theta (|den_i| < delta -> delta), r = H u - theta u.
// g = 1 is plain (diagonal) Davidson; the network
//   x_i = [asinh((diag_i - theta)/delta), (theta - mean(diag))/sd(diag),
//          log(lam_l[a] + 1e-10), log(lam_r[b] + 1e-10)],  x^ = (x - mu) * inv_sd,
// (mean and sd of the diagonal over the logical elements: theta itself grows
// with the chain length, its place in the diagonal's spread does not)
//   h1 = tanh(W1 x^ + b1), h2 = tanh(W2 h1 + b2), g = w3 . h2 + b3
// (4 -> 64 -> 64 -> 1, FP32) supplies g.  lam_l[a] / lam_r[b] are the row /
// column norms of the Ritz vector over the other two indices: the bond
// weights of the mixed-canonical state (the singular values when the gauge
// is the SVD one).  Training minimises (1/n) sum (g a - t*)^2 over samples
// a = s0 r, t* = exact correction (u0 / (u . u0) - u), on the GPU with the
// kernels below and Adam.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "tn_chain.h"

using namespace tn;

namespace tn {

constexpr int MH = 64;  // hidden width
constexpr int MLP_NP = MH * 4 + MH + MH * MH + MH + MH + 1;
constexpr int OFF_B1 = MH * 4, OFF_W2 = OFF_B1 + MH, OFF_B2 = OFF_W2 + MH * MH, OFF_W3 = OFF_B2 + MH,
              OFF_B3 = OFF_W3 + MH;

// ---------------------------------------------------------------------------
// diag[a][m][b] = sum_{w,w'} L[a][w][a] W[w][w'][m][m] R[w'][b][b]
__global__ void __launch_bounds__(256) hdiag_k(const double* __restrict__ L, const double* __restrict__ R,
                                               const double* __restrict__ Wd, int p, int D1p, int D2p, int Dw1,
                                               int Dw2, double* __restrict__ diag)
{
    const long n = (long)D1p * p * D2p;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const int b = (int)(e % D2p);
        const long am = e / D2p;
        const int m = (int)(am % p), a = (int)(am / p);
        double acc = 0.0;
        for (int w = 0; w < Dw1; w++) {
            const double l = L[((long)a * Dw1 + w) * D1p + a];
            if (l == 0.0) continue;
            double s = 0.0;
            for (int v = 0; v < Dw2; v++) s += Wd[(w * Dw2 + v) * p + m] * R[((long)v * D2p + b) * D2p + b];
            acc += l * s;
        }
        diag[e] = acc;
    }
}

// sum and sum of squares of the diagonal over the logical elements
// (a < D1, b < D2), one workgroup, into out[0], out[1]
__global__ void __launch_bounds__(256) diag_stats_k(const double* __restrict__ diag, int p, int D1, int D2, int D2p,
                                                    double* __restrict__ out)
{
    __shared__ double r1[4], r2[4];
    const long n = (long)D1 * p * D2;
    double s1 = 0.0, s2 = 0.0;
    for (long e = threadIdx.x; e < n; e += 256) {
        const long am = e / D2;
        const double v = diag[am * D2p + (e - am * D2)];
        s1 += v;
        s2 += v * v;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        r1[threadIdx.x >> 6] = s1;
        r2[threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = (r1[0] + r1[1]) + (r1[2] + r1[3]);
        out[1] = (r2[0] + r2[1]) + (r2[2] + r2[3]);
    }
}

// lam_l[a] = |u[a, :, :]|, lam_r[b] = |u[:, :, b]| (one workgroup per a / per b)
__global__ void __launch_bounds__(256) rownorm_k(const double* __restrict__ u, int p, int D2p, double* __restrict__ out)
{
    __shared__ double red[4];
    const int a = blockIdx.x;
    const long base = (long)a * p * D2p, n = (long)p * D2p;
    double s = 0.0;
    for (long i = threadIdx.x; i < n; i += 256) s += u[base + i] * u[base + i];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[a] = sqrt((red[0] + red[1]) + (red[2] + red[3]));
}
__global__ void __launch_bounds__(256) colnorm_k(const double* __restrict__ u, int D1p, int p, int D2p,
                                                 double* __restrict__ out)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= D2p) return;
    double s = 0.0;
    for (long am = 0; am < (long)D1p * p; am++) {
        const double v = u[am * D2p + b];
        s += v * v;
    }
    out[b] = sqrt(s);
}

__device__ __forceinline__ void mlp_features(double dg, double theta, double theta_rel, double ll, double lr,
                                             double delta, const float* __restrict__ nrm, float (&xh)[4])
{
    const double x0 = asinh((dg - theta) / delta), x1 = theta_rel, x2 = log(ll + 1e-10), x3 = log(lr + 1e-10);
    xh[0] = (float)((x0 - nrm[0]) * nrm[4]);
    xh[1] = (float)((x1 - nrm[1]) * nrm[5]);
    xh[2] = (float)((x2 - nrm[2]) * nrm[6]);
    xh[3] = (float)((x3 - nrm[3]) * nrm[7]);
}

// forward pass of one sample with the parameters in LDS; h1, h2 kept for the
// backward pass
__device__ __forceinline__ float mlp_fwd(const float* __restrict__ P, const float (&xh)[4], float (&h1)[MH],
                                         float (&h2)[MH])
{
#pragma unroll
    for (int j = 0; j < MH; j++) {
        float z = P[OFF_B1 + j];
#pragma unroll
        for (int k = 0; k < 4; k++) z += P[j * 4 + k] * xh[k];
        h1[j] = tanhf(z);
    }
#pragma unroll
    for (int j = 0; j < MH; j++) {
        float z = P[OFF_B2 + j];
#pragma unroll
        for (int k = 0; k < MH; k++) z += P[OFF_W2 + j * MH + k] * h1[k];
        h2[j] = tanhf(z);
    }
    float y = P[OFF_B3];
#pragma unroll
    for (int j = 0; j < MH; j++) y += P[OFF_W3 + j] * h2[j];
    return y;
}

// forward pass for inference: h2 is folded into the output sum and the
// second layer's rows are not unrolled, which keeps the live registers to h1
__device__ __forceinline__ float mlp_infer(const float* __restrict__ P, const float (&xh)[4])
{
    float h1[MH];
#pragma unroll
    for (int j = 0; j < MH; j++) {
        float z = P[OFF_B1 + j];
#pragma unroll
        for (int k = 0; k < 4; k++) z += P[j * 4 + k] * xh[k];
        h1[j] = tanhf(z);
    }
    float y = P[OFF_B3];
#pragma unroll 1
    for (int j = 0; j < MH; j++) {
        const float* w = P + OFF_W2 + j * MH;
        float z = P[OFF_B2 + j];
#pragma unroll
        for (int k = 0; k < MH; k++) z += w[k] * h1[k];
        y += P[OFF_W3 + j] * tanhf(z);
    }
    return y;
}

// t = g s0 r over the site vector; kind 1: g = 1, kind 2: g from the network
// (clamped to [-16, 16])
__global__ void __launch_bounds__(256) precond_k(const double* __restrict__ r, const double* __restrict__ diag,
                                                 double theta, double theta_rel, double delta,
                                                 const double* __restrict__ laml,
                                                 const double* __restrict__ lamr, int p, int D2p, long n, int kind,
                                                 const float* __restrict__ params, const float* __restrict__ nrm,
                                                 double* __restrict__ t)
{
    __shared__ float P[MLP_NP];
    __shared__ float Nm[8];
    if (kind == 2) {
        for (int i = threadIdx.x; i < MLP_NP; i += 256) P[i] = params[i];
        if (threadIdx.x < 8) Nm[threadIdx.x] = nrm[threadIdx.x];
        __syncthreads();
    }
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        double den = diag[e] - theta;
        if (fabs(den) < delta) den = delta;
        double g = 1.0;
        if (kind == 2) {
            const int b = (int)(e % D2p), a = (int)(e / ((long)p * D2p));
            float xh[4];
            mlp_features(diag[e], theta, theta_rel, laml[a], lamr[b], delta, Nm, xh);
            // safeguard on the magnitude of the learned factor
            g = fmin(fmax((double)mlp_infer(P, xh), -16.0), 16.0);
        }
        t[e] = -g * r[e] / den;
    }
}

// network output g for raw features X (n x 4, FP64), FP32 arithmetic
__global__ void __launch_bounds__(256) mlp_infer_k(const double* __restrict__ X, long n,
                                                   const float* __restrict__ params, const float* __restrict__ nrm,
                                                   double* __restrict__ g)
{
    __shared__ float P[MLP_NP];
    for (int i = threadIdx.x; i < MLP_NP; i += 256) P[i] = params[i];
    __syncthreads();
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        float xh[4];
#pragma unroll
        for (int k = 0; k < 4; k++) xh[k] = (float)((X[4 * e + k] - nrm[k]) * nrm[4 + k]);
        g[e] = (double)mlp_infer(P, xh);
    }
}

// Training: one workgroup = 256 samples (one per thread).  Forward and
// backward per thread; the per-sample outer products are summed over the
// block through LDS (A, B: 256 x 64), written as a partial gradient row.
constexpr int TB = 256;
__global__ void __launch_bounds__(TB) mlp_grad_k(const double* __restrict__ X, const double* __restrict__ a,
                                                 const double* __restrict__ t, const int* __restrict__ idx, long n,
                                                 const float* __restrict__ params, const float* __restrict__ nrm,
                                                 float* __restrict__ part, float* __restrict__ loss_part)
{
    __shared__ float P[MLP_NP];
    __shared__ float A[TB][MH + 1];
    __shared__ float B[TB][MH + 1];
    __shared__ float red[TB / 64];
    const int tid = threadIdx.x;
    for (int i = tid; i < MLP_NP; i += TB) P[i] = params[i];
    __syncthreads();
    const long s = (long)blockIdx.x * TB + tid;
    const bool live = s < n;
    float xh[4] = {0.f, 0.f, 0.f, 0.f}, h1[MH], h2[MH];
    float dy = 0.f, l2 = 0.f;
    {
        const long k = live ? (idx ? idx[s] : s) : 0;
#pragma unroll
        for (int q = 0; q < 4; q++) xh[q] = (float)((X[4 * k + q] - nrm[q]) * nrm[4 + q]);
        const float y = mlp_fwd(P, xh, h1, h2);
        if (live) {
            const float av = (float)a[k], res = y * av - (float)t[k];
            dy = 2.f * res * av / (float)n;
            l2 = res * res / (float)n;
        }
    }
    float* out = part + (size_t)blockIdx.x * MLP_NP;
    // d2 = dy w3 (1 - h2^2); A = d2, B = h1  -> gW2, gb2
    float d2[MH];
#pragma unroll
    for (int j = 0; j < MH; j++) {
        d2[j] = dy * P[OFF_W3 + j] * (1.f - h2[j] * h2[j]);
        A[tid][j] = d2[j];
        B[tid][j] = h1[j];
    }
    __syncthreads();
    for (int e = tid; e < MH * MH; e += TB) {
        const int j = e / MH, k = e % MH;
        float acc = 0.f;
        for (int q = 0; q < TB; q++) acc += A[q][j] * B[q][k];
        out[OFF_W2 + e] = acc;
    }
    if (tid < MH) {
        float acc = 0.f;
        for (int q = 0; q < TB; q++) acc += A[q][tid];
        out[OFF_B2 + tid] = acc;
    }
    __syncthreads();
    // d1 = (W2^T d2) (1 - h1^2); A = d1, B[:, 0..3] = x^  -> gW1, gb1
#pragma unroll
    for (int k = 0; k < MH; k++) {
        float acc = 0.f;
#pragma unroll
        for (int j = 0; j < MH; j++) acc += P[OFF_W2 + j * MH + k] * d2[j];
        A[tid][k] = acc * (1.f - h1[k] * h1[k]);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) B[tid][q] = xh[q];
    __syncthreads();
    {
        const int j = tid / 4, k = tid % 4;  // 256 = 64 x 4 entries of gW1
        float acc = 0.f;
        for (int q = 0; q < TB; q++) acc += A[q][j] * B[q][k];
        out[j * 4 + k] = acc;
    }
    if (tid < MH) {
        float acc = 0.f;
        for (int q = 0; q < TB; q++) acc += A[q][tid];
        out[OFF_B1 + tid] = acc;
    }
    __syncthreads();
    // gw3 = sum dy h2, gb3 = sum dy
#pragma unroll
    for (int j = 0; j < MH; j++) A[tid][j] = dy * h2[j];
    B[tid][0] = dy;
    __syncthreads();
    if (tid < MH) {
        float acc = 0.f;
        for (int q = 0; q < TB; q++) acc += A[q][tid];
        out[OFF_W3 + tid] = acc;
    }
    if (tid == MH) {
        float acc = 0.f;
        for (int q = 0; q < TB; q++) acc += B[q][0];
        out[OFF_B3] = acc;
    }
    // loss partial
    float lv = l2;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) lv += __shfl_xor(lv, o, 64);
    if ((tid & 63) == 0) red[tid >> 6] = lv;
    __syncthreads();
    if (tid == 0) loss_part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// grad[p] = sum over blocks of part[b][p]; loss = sum of loss_part
__global__ void __launch_bounds__(256) mlp_reduce_k(const float* __restrict__ part, const float* __restrict__ lp,
                                                    int nb, float* __restrict__ grad, float* __restrict__ loss)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < MLP_NP) {
        float acc = 0.f;
        for (int b = 0; b < nb; b++) acc += part[(size_t)b * MLP_NP + p];
        grad[p] = acc;
    }
    if (p == 0) {
        float acc = 0.f;
        for (int b = 0; b < nb; b++) acc += lp[b];
        loss[0] = acc;
    }
}

// Adam step (Kingma & Ba 2015), bias-corrected with the step number
__global__ void __launch_bounds__(256) adam_k(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, float lr, float b1, float b2, float eps,
                                              float c1, float c2)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= MLP_NP) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp, vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// ---------------------------------------------------------------------------
// host helpers
void heff_diag(tn_ctx* c, const HeffOp& op, const double* Wd, double* diag)
{
    const long n = op.n();
    long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hdiag_k<<<blocks, 256, 0, c->stream>>>(op.L, op.R, Wd, op.p, op.D1p, op.D2p, op.Dw1, op.Dw2, diag);
    TN_HIP(hipGetLastError());
}

void bond_weights(tn_ctx* c, const double* u, int p, int D1p, int D2p, double* laml, double* lamr)
{
    rownorm_k<<<D1p, 256, 0, c->stream>>>(u, p, D2p, laml);
    colnorm_k<<<(D2p + 255) / 256, 256, 0, c->stream>>>(u, D1p, p, D2p, lamr);
    TN_HIP(hipGetLastError());
}

void precond_apply(tn_ctx* c, int kind, const double* r, const double* diag, double theta, double theta_rel,
                   double delta, const double* laml, const double* lamr, int p, int D2p, long n, double* t)
{
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    const float* params = c->mlp.as<float>();
    precond_k<<<blocks, 256, 0, c->stream>>>(r, diag, theta, theta_rel, delta, laml, lamr, p, D2p, n, kind, params,
                                             params ? params + MLP_NP : nullptr, t);
    TN_HIP(hipGetLastError());
}

// lowest two eigenpairs of a small symmetric matrix (cyclic Jacobi), ascending
void sym_eig_small(int m, const std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z)
{
    std::vector<double> a = A;
    Z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++) Z[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                tot += a[(size_t)i * m + j] * a[(size_t)i * m + j];
                if (i != j) off += a[(size_t)i * m + j] * a[(size_t)i * m + j];
            }
        if (off <= 1e-30 * tot) break;
        for (int pp = 0; pp < m - 1; pp++)
            for (int q = pp + 1; q < m; q++) {
                const double apq = a[(size_t)pp * m + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = a[(size_t)pp * m + pp], aqq = a[(size_t)q * m + q];
                const double tau = (aqq - app) / (2.0 * apq);
                const double tt = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double cs = 1.0 / std::sqrt(1.0 + tt * tt), sn = tt * cs;
                for (int k = 0; k < m; k++) {
                    const double akp = a[(size_t)k * m + pp], akq = a[(size_t)k * m + q];
                    a[(size_t)k * m + pp] = cs * akp - sn * akq;
                    a[(size_t)k * m + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < m; k++) {
                    const double apk = a[(size_t)pp * m + k], aqk = a[(size_t)q * m + k];
                    a[(size_t)pp * m + k] = cs * apk - sn * aqk;
                    a[(size_t)q * m + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < m; k++) {
                    const double zkp = Z[(size_t)k * m + pp], zkq = Z[(size_t)k * m + q];
                    Z[(size_t)k * m + pp] = cs * zkp - sn * zkq;
                    Z[(size_t)k * m + q] = sn * zkp + cs * zkq;
                }
            }
    }
    std::vector<int> ord(m);
    for (int i = 0; i < m; i++) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a[(size_t)x * m + x] < a[(size_t)y * m + y]; });
    w.resize(m);
    std::vector<double> Zs((size_t)m * m);
    for (int j = 0; j < m; j++) {
        w[j] = a[(size_t)ord[j] * m + ord[j]];
        for (int k = 0; k < m; k++) Zs[(size_t)k * m + j] = Z[(size_t)k * m + ord[j]];
    }
    Z.swap(Zs);
}

}  // namespace tn

// ---------------------------------------------------------------------------
// Davidson on the site problem: V (basis), W = H V, projected matrix on the
// host.  kind 0: no preconditioner (t = r), 1: diagonal, 2: network.  The
// correction is orthogonalised against the basis by two classical
// Gram-Schmidt passes.  Restart (thick, to the Ritz vector) at mmax vectors.
namespace tn {

DavidsonResult davidson(tn_ctx* c, ApplyFn fn, const void* op_, const HeffOp& geo, int D1, int D2, const double* Wd,
                        double* x, double tol, int mmax, int max_iter, int kind, double delta)
{
    hipStream_t s = c->stream;
    DavidsonResult res;
    const long N = geo.n();
    mmax = std::max(3, std::min(mmax, 200));
    c->krylov.ensure((size_t)mmax * N * 8);
    c->dav_W.ensure((size_t)mmax * N * 8);
    c->dav_vec.ensure((size_t)4 * N * 8 + (size_t)(geo.D1p + geo.D2p) * 8 + 4096);
    double* V = c->krylov.as<double>();
    double* Wb = c->dav_W.as<double>();
    double* u = c->dav_vec.as<double>();
    double* Hu = u + N;
    double* diag = Hu + N;
    double* rt = diag + N;  // residual
    double* laml = rt + N;
    double* lamr = laml + geo.D1p;
    c->lanczos_small.ensure(sizeof(double) * (4 * 512 + 64));
    double* h = c->lanczos_small.as<double>();
    const long pcap = mdot_part_size(N, mmax + 1);
    c->part.ensure(pcap * 8);
    double* part = c->part.as<double>();
    if (kind >= 2) TN_REQUIRE(c->mlp.p != nullptr, TN_ERR_STATE, "Davidson: no preconditioner network loaded");
    double dmu = 0.0, dsd = 1.0;
    if (kind >= 1) {
        heff_diag(c, geo, Wd, diag);
        if (kind >= 2) {
            double st[2];
            diag_stats_k<<<1, 256, 0, s>>>(diag, geo.p, D1, D2, geo.D2p, h);
            TN_HIP(hipGetLastError());
            d2h(c, st, h, 16);
            const double nl = (double)D1 * geo.p * D2;
            dmu = st[0] / nl;
            dsd = std::sqrt(std::max(st[1] / nl - dmu * dmu, 0.0));
            if (!(dsd > 1e-12)) dsd = 1.0;
        }
    }
    auto dot_host = [&](const double* a, long lda, int k, const double* w, double* out) {
        TN_HIP(mdot(a, lda, k, w, N, h, part, (int)pcap, s));
        d2h(c, out, h, (size_t)k * 8);
    };
    double nx = 0;
    dot_host(x, N, 1, x, &nx);
    nx = std::sqrt(nx);
    TN_REQUIRE(nx > 0 && std::isfinite(nx), TN_ERR_STATE, "Davidson: zero start vector");
    TN_HIP(scale_copy(1.0 / nx, x, V, N, s));
    fn(c, op_, V, Wb);
    res.n_matvec++;
    std::vector<double> Hm((size_t)mmax * mmax, 0.0), hh(mmax + 2), w, Z;
    dot_host(V, N, 1, Wb, hh.data());
    Hm[0] = hh[0];
    int m = 1;
    for (int it = 0; it < max_iter; it++) {
        std::vector<double> A((size_t)m * m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) A[(size_t)i * m + j] = Hm[(size_t)i * mmax + j];
        sym_eig_small(m, A, w, Z);
        res.e0 = w[0];
        res.e1 = m > 1 ? w[1] : w[0];
        std::vector<double> y(m);
        for (int i = 0; i < m; i++) y[i] = Z[(size_t)i * m];
        h2d(c, h, y.data(), (size_t)m * 8);
        TN_HIP(gemv_acc(V, N, m, h, 1.0, 0.0, u, N, s));
        TN_HIP(gemv_acc(Wb, N, m, h, 1.0, 0.0, Hu, N, s));
        TN_HIP(scale_copy(1.0, Hu, rt, N, s));
        TN_HIP(axpby(-res.e0, u, 1.0, rt, N, s));
        double r2 = 0;
        dot_host(rt, N, 1, rt, &r2);
        res.residual = std::sqrt(std::max(r2, 0.0));
        res.iters = it;
        if (res.residual < tol) {
            res.converged = true;
            break;
        }
        if (m == mmax) {  // restart from the Ritz pair
            TN_HIP(hipMemcpyAsync(V, u, N * 8, hipMemcpyDeviceToDevice, s));
            TN_HIP(hipMemcpyAsync(Wb, Hu, N * 8, hipMemcpyDeviceToDevice, s));
            Hm.assign((size_t)mmax * mmax, 0.0);
            Hm[0] = res.e0;
            m = 1;
        }
        double* t = V + (long)m * N;
        // kind 3: the network on the first correction from the start vector
        // (the step its training samples come from), the diagonal one after
        const int ks = kind == 3 ? (it == 0 ? 2 : 1) : kind;
        if (ks == 0) {
            TN_HIP(scale_copy(1.0, rt, t, N, s));
        } else {
            if (ks == 2) bond_weights(c, u, geo.p, geo.D1p, geo.D2p, laml, lamr);
            precond_apply(c, ks, rt, diag, res.e0, (res.e0 - dmu) / dsd, delta, laml, lamr, geo.p, geo.D2p, N, t);
        }
        for (int pass = 0; pass < 2; pass++) {
            TN_HIP(mdot(V, N, m, t, N, h, part, (int)pcap, s));
            TN_HIP(gemv_acc(V, N, m, h, -1.0, 1.0, t, N, s));
        }
        double nt = 0;
        dot_host(t, N, 1, t, &nt);
        nt = std::sqrt(std::max(nt, 0.0));
        if (!(nt > 1e-14)) break;  // the correction lies in the basis: stagnation
        TN_HIP(scale_copy(1.0 / nt, t, t, N, s));
        fn(c, op_, t, Wb + (long)m * N);
        res.n_matvec++;
        dot_host(V, N, m + 1, Wb + (long)m * N, hh.data());
        for (int i = 0; i <= m; i++) Hm[(size_t)i * mmax + m] = Hm[(size_t)m * mmax + i] = hh[i];
        m++;
    }
    // x = u (normalised Ritz vector)
    double nu = 0;
    dot_host(u, N, 1, u, &nu);
    TN_HIP(scale_copy(1.0 / std::sqrt(nu), u, x, N, s));
    sync(c);
    return res;
}

}  // namespace tn

// ---------------------------------------------------------------------------
extern "C" {

namespace {
int fail_ctx(tn_ctx* c, const TnError& e)
{
    if (c) c->err = e.msg;
    return e.code;
}
}  // namespace

int tn_ctx_set_mlp(tn_ctx* c, const float* params, const float* norm)
{
    if (!c || !params || !norm) return TN_ERR_ARG;
    try {
        c->mlp.ensure((MLP_NP + 8) * sizeof(float));
        h2d(c, c->mlp.p, params, MLP_NP * sizeof(float));
        h2d(c, c->mlp.as<float>() + MLP_NP, norm, 8 * sizeof(float));
        return TN_OK;
    } catch (const TnError& e) {
        return fail_ctx(c, e);
    }
}

int tn_mlp_forward(tn_ctx* c, const float* params, const float* norm, int64_t n, const double* X, double* g)
{
    if (!c || !params || !norm || n < 0 || (n > 0 && (!X || !g))) return TN_ERR_ARG;
    if (n == 0) return TN_OK;
    try {
        DBuf p, x, y;
        p.ensure((MLP_NP + 8) * 4);
        x.ensure((size_t)n * 4 * 8);
        y.ensure((size_t)n * 8);
        h2d(c, p.p, params, MLP_NP * 4);
        h2d(c, p.as<float>() + MLP_NP, norm, 32);
        h2d(c, x.p, X, (size_t)n * 32);
        long blocks = (n + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        mlp_infer_k<<<blocks, 256, 0, c->stream>>>(x.as<double>(), n, p.as<float>(), p.as<float>() + MLP_NP,
                                                   y.as<double>());
        TN_HIP(hipGetLastError());
        d2h(c, g, y.p, (size_t)n * 8);
        return TN_OK;
    } catch (const TnError& e) {
        return fail_ctx(c, e);
    }
}

namespace {
// gradient of the mean loss over the samples idx[0..n) (or all when idx is
// null) into grad (device), loss into loss (device); parameters on the device
void mlp_grad_dev(tn_ctx* c, const float* P, const double* X, const double* a, const double* t, const int* idx,
                  long n, DBuf& part, DBuf& lpart, float* grad, float* loss)
{
    const int nb = (int)((n + TB - 1) / TB);
    part.ensure((size_t)nb * MLP_NP * 4);
    lpart.ensure((size_t)nb * 4 + 256);
    mlp_grad_k<<<nb, TB, 0, c->stream>>>(X, a, t, idx, n, P, P + MLP_NP, part.as<float>(), lpart.as<float>());
    mlp_reduce_k<<<(MLP_NP + 255) / 256, 256, 0, c->stream>>>(part.as<float>(), lpart.as<float>(), nb, grad, loss);
    TN_HIP(hipGetLastError());
}
}  // namespace

int tn_mlp_loss_grad(tn_ctx* c, const float* params, const float* norm, int64_t n, const double* X,
                     const double* a, const double* t, double* loss, float* grad)
{
    if (!c || !params || !norm || n <= 0 || !X || !a || !t || !grad) return TN_ERR_ARG;
    try {
        DBuf p, x, av, tv, part, lp, g;
        p.ensure((MLP_NP + 8) * 4);
        x.ensure((size_t)n * 32);
        av.ensure((size_t)n * 8);
        tv.ensure((size_t)n * 8);
        g.ensure((MLP_NP + 8) * 4);
        h2d(c, p.p, params, MLP_NP * 4);
        h2d(c, p.as<float>() + MLP_NP, norm, 32);
        h2d(c, x.p, X, (size_t)n * 32);
        h2d(c, av.p, a, (size_t)n * 8);
        h2d(c, tv.p, t, (size_t)n * 8);
        mlp_grad_dev(c, p.as<float>(), x.as<double>(), av.as<double>(), tv.as<double>(), nullptr, n, part, lp,
                     g.as<float>(), g.as<float>() + MLP_NP);
        std::vector<float> h(MLP_NP + 1);
        d2h(c, h.data(), g.p, (MLP_NP + 1) * 4);
        std::memcpy(grad, h.data(), MLP_NP * 4);
        if (loss) *loss = h[MLP_NP];
        return TN_OK;
    } catch (const TnError& e) {
        return fail_ctx(c, e);
    }
}

int tn_mlp_train(tn_ctx* c, float* params, const float* norm, int64_t n, const double* X, const double* a,
                 const double* t, int steps, int batch, double lr, uint64_t seed, double* loss)
{
    if (!c || !params || !norm || n <= 0 || !X || !a || !t || steps < 0 || batch <= 0) return TN_ERR_ARG;
    try {
        hipStream_t s = c->stream;
        DBuf p, x, av, tv, part, lp, g, m, v, idx;
        p.ensure((MLP_NP + 8) * 4);
        x.ensure((size_t)n * 32);
        av.ensure((size_t)n * 8);
        tv.ensure((size_t)n * 8);
        g.ensure((MLP_NP + 8) * 4);
        m.ensure(MLP_NP * 4);
        v.ensure(MLP_NP * 4);
        const int B = (int)std::min<int64_t>(batch, n);
        idx.ensure((size_t)steps * B * 4 + 256);
        h2d(c, p.p, params, MLP_NP * 4);
        h2d(c, p.as<float>() + MLP_NP, norm, 32);
        h2d(c, x.p, X, (size_t)n * 32);
        h2d(c, av.p, a, (size_t)n * 8);
        h2d(c, tv.p, t, (size_t)n * 8);
        TN_HIP(fill_zero(m.p, MLP_NP * 4, s));
        TN_HIP(fill_zero(v.p, MLP_NP * 4, s));
        // minibatch indices: seeded draws without replacement within a batch
        std::vector<int> ids((size_t)steps * B);
        std::mt19937_64 rng(seed);
        std::vector<int> perm(n);
        for (int64_t k = 0; k < n; k++) perm[k] = (int)k;
        size_t pos = n;
        for (size_t k = 0; k < ids.size(); k++) {
            if (pos >= (size_t)n) {
                std::shuffle(perm.begin(), perm.end(), rng);
                pos = 0;
            }
            ids[k] = perm[pos++];
        }
        if (!ids.empty()) h2d(c, idx.p, ids.data(), ids.size() * 4);
        float* P = p.as<float>();
        float* G = g.as<float>();
        auto full_loss = [&]() {
            mlp_grad_dev(c, P, x.as<double>(), av.as<double>(), tv.as<double>(), nullptr, n, part, lp, G,
                         G + MLP_NP);
            float l = 0;
            d2h(c, &l, G + MLP_NP, 4);
            return (double)l;
        };
        if (loss) loss[0] = full_loss();
        const float b1 = 0.9f, b2 = 0.999f;
        for (int k = 0; k < steps; k++) {
            mlp_grad_dev(c, P, x.as<double>(), av.as<double>(), tv.as<double>(), idx.as<int>() + (size_t)k * B, B,
                         part, lp, G, G + MLP_NP);
            const float c1 = 1.f - std::pow(b1, (float)(k + 1)), c2 = 1.f - std::pow(b2, (float)(k + 1));
            adam_k<<<(MLP_NP + 255) / 256, 256, 0, s>>>(P, G, m.as<float>(), v.as<float>(), (float)lr, b1, b2, 1e-8f,
                                                        c1, c2);
            TN_HIP(hipGetLastError());
        }
        if (loss) loss[1] = full_loss();
        d2h(c, params, P, MLP_NP * 4);
        return TN_OK;
    } catch (const TnError& e) {
        return fail_ctx(c, e);
    }
}

int tn_chain_precond_samples(tn_chain* ch, int site, int iters, int nsamp, uint64_t seed, double delta, double* X,
                             double* a, double* t, int* nout)
{
    if (!ch || site < 0 || site >= ch->L || iters <= 0 || nsamp <= 0 || !X || !a || !t || !nout) return TN_ERR_ARG;
    tn_ctx* c = ch->ctx;
    try {
        hipStream_t s = c->stream;
        HeffCall hc;
        hc.op = make_heff(ch, site, TN_FP64, 0);
        hc.st = nullptr;
        hc.D1 = ch->D1[site];
        hc.D2 = ch->D2[site];
        const HeffOp& op = hc.op;
        const long N = op.n();
        const int p = op.p, D1p = op.D1p, D2p = op.D2p;
        DBuf u0, x0, diag, uu, rr, ll, lr;
        for (auto* b : {&u0, &x0, &diag, &uu, &rr}) b->ensure((size_t)N * 8);
        ll.ensure((size_t)D1p * 8);
        lr.ensure((size_t)D2p * 8);
        // exact local ground state (tight Lanczos from the site tensor)
        TN_HIP(hipMemcpyAsync(u0.p, ch->A[site], N * 8, hipMemcpyDeviceToDevice, s));
        LanczosResult lz = lanczos(c, heff_cb, &hc, N, u0.as<double>(), 1e-12, 120, 20);
        (void)lz;
        TN_HIP(hipMemcpyAsync(x0.p, ch->A[site], N * 8, hipMemcpyDeviceToDevice, s));
        const double* Wd = w_diag(ch, site);
        heff_diag(c, op, Wd, diag.as<double>());
        std::vector<double> hu0(N), hdiag(N), hu(N), hr(N), hll(D1p), hlr(D2p);
        d2h(c, hu0.data(), u0.p, N * 8);
        d2h(c, hdiag.data(), diag.p, N * 8);
        std::mt19937_64 rng(seed);
        // logical elements only (padding stays zero)
        std::vector<long> logical;
        for (int aa = 0; aa < ch->D1[site]; aa++)
            for (int m = 0; m < p; m++)
                for (int b = 0; b < ch->D2[site]; b++) logical.push_back(((long)aa * p + m) * D2p + b);
        double dmu = 0.0, dsq = 0.0;
        for (long e : logical) {
            dmu += hdiag[e];
            dsq += hdiag[e] * hdiag[e];
        }
        dmu /= (double)logical.size();
        double dsd = std::sqrt(std::max(dsq / (double)logical.size() - dmu * dmu, 0.0));
        if (!(dsd > 1e-12)) dsd = 1.0;
        int w = 0;
        // Davidson with the diagonal correction from the site tensor; after
        // k iterations (k = 1..iters) the Ritz pair is read back and sampled
        for (int k = 1; k <= iters; k++) {
            TN_HIP(hipMemcpyAsync(uu.p, x0.p, N * 8, hipMemcpyDeviceToDevice, s));
            DavidsonResult dr = davidson(c, heff_cb, &hc, op, ch->D1[site], ch->D2[site], Wd, uu.as<double>(), 0.0,
                                         64, k, 1, delta);
            // residual of the returned Ritz vector
            heff_cb(c, &hc, uu.as<double>(), rr.as<double>());
            TN_HIP(axpby(-dr.e0, uu.as<double>(), 1.0, rr.as<double>(), N, s));
            bond_weights(c, uu.as<double>(), p, D1p, D2p, ll.as<double>(), lr.as<double>());
            d2h(c, hu.data(), uu.p, N * 8);
            d2h(c, hr.data(), rr.p, N * 8);
            d2h(c, hll.data(), ll.p, D1p * 8);
            d2h(c, hlr.data(), lr.p, D2p * 8);
            double uu0 = 0, nr = 0;
            for (long e : logical) {
                uu0 += hu[e] * hu0[e];
                nr += hr[e] * hr[e];
            }
            nr = std::sqrt(nr);
            if (!(nr > 0) || std::fabs(uu0) < 1e-12) break;
            std::uniform_int_distribution<size_t> pick(0, logical.size() - 1);
            for (int q = 0; q < nsamp; q++) {
                const long e = logical[pick(rng)];
                const int b = (int)(e % D2p), aa = (int)(e / ((long)p * D2p));
                double den = hdiag[e] - dr.e0;
                if (std::fabs(den) < delta) den = delta;
                X[4 * w + 0] = std::asinh((hdiag[e] - dr.e0) / delta);
                X[4 * w + 1] = (dr.e0 - dmu) / dsd;
                X[4 * w + 2] = std::log(hll[aa] + 1e-10);
                X[4 * w + 3] = std::log(hlr[b] + 1e-10);
                a[w] = -hr[e] / den / nr;
                t[w] = (hu0[e] / uu0 - hu[e]) / nr;
                w++;
            }
        }
        *nout = w;
        sync(c);
        return TN_OK;
    } catch (const TnError& e) {
        return fail_ctx(c, e);
    }
}

}  // extern "C"
