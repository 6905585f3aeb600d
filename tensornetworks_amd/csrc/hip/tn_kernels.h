// tn_kernels.h -- launch wrappers for the gfx950 kernels of the DMRG hot path.
// Internal header (not part of the C ABI).  All launches are asynchronous on
// the given stream; pointers are device pointers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tn {

typedef uint16_t bf16;  // storage type of bfloat16 operands

enum Prec { P_FP64 = 0, P_FP32 = 1, P_BF16 = 2 };
// T_BF16X3: a value held as three bfloat16 planes v = v0 + v1 + v2 (each the
// rounding of what the previous ones leave, about 24 significant bits
// together); products of such operands sum the six leading cross products
// A_i B_j (i + j <= 2) on the bfloat16 matrix cores in FP32.  dtype_size is
// the size of one plane's element.
enum DType { T_F64 = 0, T_F32 = 1, T_BF16 = 2, T_BF16X3 = 3 };

inline size_t dtype_size(int t) { return t == T_F64 ? 8 : (t == T_F32 ? 4 : 2); }

// C[b] = alpha * op(A[b]) * op(B[b]) (+ C[b] if accumulate)
//   a_kmajor: A(m,k) = A[m*lda + k]   else A[k*lda + m]
//   b_kmajor: B(k,n) = B[n*ldb + k]   else B[k*ldb + n]
// M, N must be multiples of 32, K a multiple of 32 (operands are zero padded).
struct GemmDesc {
    int tin;   // DType of A and B
    int tout;  // DType of C
    bool a_kmajor, b_kmajor;
    int M, N, K, batch;
    const void* A; long lda, sA;
    const void* B; long ldb, sB;
    void* C; long ldc, sC;
    float alpha;
    int accumulate;
    int tile;  // 0 = automatic, else forced tile code (autotuner)
    // T_BF16X3 operands: plane p of A starts at A + p*psA elements (likewise
    // B); a_kmajor && b_kmajor only
    long psA = 0, psB = 0;
};
hipError_t gemm(const GemmDesc& g, hipStream_t s);

// One problem of a grouped GEMM (device-resident descriptor): C = alpha A B
// (+ C), layouts and types shared by the whole group.
struct GemmProb {
    const void* A;
    const void* B;
    void* C;
    long lda, ldb, ldc;
    int M, N, K;
    float alpha;
    int accumulate;
};
// One persistent launch over all output tiles of nprob problems; off (device,
// nprob+1) is the prefix sum of per-problem tile counts (M/BM)*(N/BN) for the
// tile code; bk the K step (gemm_k_step).  M, N of every problem must be
// multiples of the tile, K of bk.
hipError_t gemm_grouped(int tin, int tout, bool akm, bool bkm, int tile, int bk, const GemmProb* probs,
                        const int* off, int nprob, int total_tiles, hipStream_t s);
int gemm_tile_bm(int code);
// tile code usable for an M x N x K problem of input type tin
bool gemm_tile_ok(int code, int tin, int M, int N, int K);
int gemm_tile_bn(int code);
int gemm_default_tile(int M, int N, int batch);
int gemm_k_step(int tin, int K);
// split-K epilogue: C[i*ldc + j] = (accumulate ? C : 0) + sum_{z<ks} W[z][i][j]
// (W: ks x M x N of type tacc, T_F64 or T_F32)
hipError_t splitk_reduce(int tacc, const void* W, int ks, int M, int N, int tout, void* C, long ldc, int accumulate,
                         hipStream_t s);
int gemm_num_tiles();
const char* gemm_tile_name(int code);

// MPO mixing: for every position (x, b), out[x*osx + ooff[o] + b] =
//   sum_{e in optr[o]..optr[o+1]} eval[e] * in[x*isx + eoff[e] + b]
struct MixTable {
    int nout, nnz;
    const int* optr;     // device, nout+1
    const long* ooff;    // device, nout
    const long* eoff;    // device, nnz
    const double* eval;  // device, nnz
    const float* evalf;  // device, nnz (same values, fp32)
    const int* esel;     // device, nnz (identity-channel tables only): 1 = entry reads in2
    const int* osel;     // device, nout (identity-channel tables only): 1 = output goes to out2
    // MIX_HEFF_T only, for the fused stage 1 + MPO stage: per output
    // m | (kept left channel << 8) | (osel << 16), per entry
    // n | (kept right channel << 8) | (esel << 16)
    const int* ocode;
    const int* ecode;
};
hipError_t wmix(int tin, int tout, const void* in, long isx, void* out, long osx, int X,
                int Bn, const MixTable& t, hipStream_t s);
// Stage 1 of the split-bfloat16 Heff apply with the MPO stage in its
// epilogue: T1 = X R^T tile by tile (X: planes of x, rows (a', n), K-major,
// ld D2p; R: planes of the nR kept right channels, rows (w', b), ld D2p), then
// in LDS T2t[(m, b)][(w, a')] = sum W T1 (+ the right identity channel from
// the FP64 x), written as three bfloat16 planes (plane stride ps2, row
// stride ldo), and the left identity channel's slice straight into y[a'][m][b]
// (FP64).  One workgroup per 32 a' x 32 b.  Returns hipErrorInvalidValue for
// shapes it does not cover (p not 2 or 3, nR not 2..5, tables larger than
// 32 outputs / 256 entries): the caller then runs the two-kernel path.
bool x3_stage1_mix_ok(int p, int nR, const MixTable& t, int D1p, int D2p);
hipError_t x3_stage1_mix(int p, int nR, const bf16* X, long psX, const bf16* R, long psR, const double* x,
                         bf16* T2, long ps2, long ldo, double* y, int D1p, int D2p, const MixTable& t,
                         hipStream_t s);
// The Heff mixing with a second input and a second output, its stage-3
// operand written transposed: entries e with esel[e] read
// in2[x*isx2 + eoff[e] + b] (type tx), outputs o with osel[o] write
// out2[x*osx2 + ooff[o] + b] (fp64).  They serve an environment channel that
// is the identity: its stage-1 slice is the input vector itself (in2) and its
// stage-3 contribution is a copy of the mixed slice into y (out2).
// Outputs o with osel[o] = 0 go to out[ooff[o] + b*ldo + x] (rows (m, b),
// columns (w, a')),
// through an LDS tile of 32 x-values by 64 b-values so both the reads (along
// b) and the writes (along x) are coalesced.  X must be a multiple of 32.
// tout = T_BF16X3: out is written as three bfloat16 planes ops elements apart
// (with tin = T_F32 and the second input tx = T_F64)
hipError_t wmix_idt(int tin, int tx, int tout, const void* in, long isx, const void* in2, long isx2, void* out,
                    long ldo, double* out2, long osx2, int X, int Bn, const MixTable& t, hipStream_t s,
                    long ops = 0);

// Small device -> host transfer without a stream synchronisation: copies n
// (<= TN_PUBLISH_MAX) doubles to host-mapped coherent memory, then
// (system-scope fence) stores seq into *flag; the host polls the flag.
#define TN_PUBLISH_MAX 16384
hipError_t publish(const double* src, int n, double* host_dst, unsigned long long* host_flag,
                   unsigned long long seq, hipStream_t s);

// elementwise
hipError_t convert(int tin, const void* in, int tout, void* out, long n, hipStream_t s);
// three bfloat16 planes of n values (T_F64 or T_F32 input): out[p*ps + i], p = 0..2
hipError_t split3(int tin, const void* in, void* out, long n, long ps, hipStream_t s);
// x <- a x in place (FP64) and the scaled values as three bfloat16 planes /
// converted to tout (T_F32 or T_BF16)
hipError_t split3_scale(double* x, double a, void* out, long n, long ps, hipStream_t s);
hipError_t convert_scale(double* x, double a, int tout, void* out, long n, hipStream_t s);
hipError_t fill_zero(void* p, long bytes, hipStream_t s);
hipError_t fill_value(double* p, double v, long n, hipStream_t s);

// site tensor layouts: reference psi[m + d*(a + D1*b)] <-> internal
// out[(a*d + m)*D2p + b] over [D1p][d][D2p] (padding written as zero)
hipError_t psi_ref_to_int(const double* ref, double* out, int d, int D1, int D2, int D1p, int D2p, hipStream_t s);
hipError_t psi_int_to_ref(const double* in, double* ref, int d, int D1, int D2, int D2p, hipStream_t s);

// copy a 3-index tensor with a strided view: out[(i*n1 + j)*n2o + k] = in[i*si + j*sj + k*sk]
// for i<n0, j<n1, k<n2; entries k in [n2, n2o) are zeroed.
hipError_t copy3(const double* in, long si, long sj, long sk, double* out, int n0, int n1,
                 int n2, int n2o, hipStream_t s);
// general strided 2-D copy with zero fill of the destination block
// out[i*so0 + j*so1] = (i<r && j<c) ? in[i*si0 + j*si1] : 0, for i<ro, j<co
hipError_t copy2(const double* in, long si0, long si1, int r, int c, double* out, long so0,
                 long so1, int ro, int co, hipStream_t s);

// Lanczos vector kernels (fp64)
// out[i] = sum_n V[i*ldv + n] * w[n], i < k     (two-stage deterministic)
hipError_t mdot(const double* V, long ldv, int k, const double* w, long n, double* out,
                double* part, int part_cap, hipStream_t s);
long mdot_part_size(long n, int k);
// w[n] = beta_w * w[n] + sum_i coef[i] * V[i*ldv + n]
hipError_t gemv_acc(const double* V, long ldv, int k, const double* coef, double coef_scale,
                    double beta_w, double* w, long n, hipStream_t s);
// out[0] = max_i |x[i]|  (part: >= 1024 doubles of scratch)
hipError_t absmax(const double* x, long n, double* out, double* part, int part_cap, hipStream_t s);
// v[n] *= 1/den[0] (den on device)
hipError_t scale_by_inv(double* v, const double* den, long n, hipStream_t s);
hipError_t axpby(double a, const double* x, double b, double* y, long n, hipStream_t s);
// w -= alpha[0] v + beta u (u may be null); alpha is a device scalar
// Fused Lanczos step (n even, vectors 16-byte aligned, k <= 512):
//  step_dots : w <- w - alpha[0] V_{k-1} - beta V_{k-2} (the latter when k >= 2),
//              then out[i] = V_i . w for i < k (two-stage, fixed order)
//  update_norm: w <- w - sum_{i<k} coef[i] V_i, then out[0] = |w|^2 (two-stage)
// part: mdot_part_size(n, k + 1) doubles.
hipError_t lz_step_dots(double* w, const double* V, long ldv, int k, const double* alpha, double beta, long n,
                        double* out, double* part, long part_cap, hipStream_t s);
hipError_t lz_update_norm(double* w, const double* V, long ldv, int k, const double* coef, long n, double* out,
                          double* part, long part_cap, hipStream_t s);
hipError_t three_term(double* w, const double* v, const double* u, const double* alpha, double beta, long n,
                      hipStream_t s);
// y = a x (y is not read; x and y may be the same buffer)
hipError_t scale_copy(double a, const double* x, double* y, long n, hipStream_t s);

// alpha[j] = h[j]   (copy a device scalar)
hipError_t dcopy_scalar(const double* src, double* dst, hipStream_t s);
// dst[0] = sqrt(src[0])
hipError_t dsqrt_scalar(const double* src, double* dst, hipStream_t s);

// one-sided block Jacobi SVD round (fp64, column-major iterate with ldx,
// q%32==0): the round's pairs rotate Xs into Xd (every column of the round's
// pairs written, rotated or copied; Xs != Xd) and V in place, nchunk
// workgroups per pair, each its own chunk of rows; each pair's 32 x 32 Gram
// matrix gets at most `inner` sweeps of rotations
hipError_t jacobi_step(const double* Xs, double* Xd, long ldx, int p, double* V, long ldv, int qv,
                       const int* pairs, int npairs, int nchunk, double tol, int* flag, hipStream_t s, int inner);
// Householder QR of the 32-column panel c0..c0+31 of the column-major A (pp
// rows, ld lda), in place: R on and above the diagonal, zeros below; the
// vectors in columns 0..31 of V (rows c0..pp-1, ld ldv) and the 32 x 32
// upper-triangular T (row-major) of the panel's block reflector I - V T V^T.
hipError_t hqr_panel(double* A, long lda, int pp, int c0, double* V, long ldv, double* T, hipStream_t s);
// m[i] = max_{j != i, j < n} |G_ij| / sqrt(G_ii G_jj) (G row-major, ld ldg; pairs
// with G_ii G_jj <= 1e-300 skipped): the largest cosine between columns
hipError_t max_cos(const double* G, long ldg, int n, double* m, hipStream_t s);
// out[j*ldo + i] = in[perm[j]*ldi + i], i < rows, j < ncols (perm: device, < the input's columns)
hipError_t gather_cols(const double* in, long ldi, double* out, long ldo, const int* perm, int rows, int ncols,
                       hipStream_t s);
// X[j*ldx + perm[i]] = R[i*ldr + j], i, j < n (perm: device, a permutation of 0..n-1)
hipError_t tr_scatter(const double* R, long ldr, double* X, long ldx, const int* perm, int n, hipStream_t s);
hipError_t col_norms(const double* X, long ldx, int p, int q, double* s, hipStream_t st);
hipError_t set_int(int* p, int v, hipStream_t s);
hipError_t set_identity(double* V, long ldv, int n, hipStream_t s);
// out1[i*s10 + j*s11] = X[perm[j]*ldx + i] * scl[j]   (i < p, j < k)
// out2[j*s20 + i*s21] = V[perm[j]*ldv + i]            (i < q, j < k)
hipError_t svd_gather(const double* X, long ldx, int p, const double* V, long ldv, int q,
                      const int* perm, const double* scl, int k, double* out1, long s10, long s11,
                      double* out2, long s20, long s21, hipStream_t s);
// unit vector e_pos in column j of X (length p)
hipError_t set_unit_col(double* x, int p, int pos, hipStream_t s);

// scale columns / rows: A[i*lda + j] *= sv[j] (cols) or sv[i] (rows), i<r, j<c
hipError_t scale_cols(double* A, long lda, int r, int c, const double* sv, hipStream_t s);
hipError_t scale_rows(double* A, long lda, int r, int c, const double* sv, hipStream_t s);

// CholeskyQR pieces (fp64): chol_diag factors the nbk x nbk (nbk <= 64)
// diagonal block at k0 of G into L (lower, ldl) and writes its inverse into
// the same block of Linv (ldi); flag[0]=1 on a non-positive pivot.
// shift_diag: G[i][i] += factor * trace(G), i < n.  add_diag: G[i][i] += v,
// i in [i0,i1).  eye_deviation: out[0] = max |G - I| over the leading n x n
// block (part: 1024 doubles).
hipError_t chol_diag(const double* G, long ldg, int k0, int nbk, double* L, long ldl, double* Linv, long ldi,
                     int* flag, hipStream_t s);
hipError_t shift_diag(double* G, long ldg, int n, double factor, hipStream_t s);
// E = G - I on the leading nl x nl block of an n x n matrix, zero elsewhere
hipError_t eye_sub(const double* G, int n, int nl, double* E, hipStream_t s);
// X = I + c1 E + c2 E2 + c3 E3, Y = I + d1 E + d2 E2 + d3 E3 (n x n)
hipError_t series3(const double* E, const double* E2, const double* E3, int n, const double c[3], const double d[3],
                   double* X, double* Y, hipStream_t s);

// --- batched (one job per chain) kernels ---------------------------------
struct CholJob {
    const double* G;
    long ldg;
    int k0, nbk;
    double* L;
    long ldl;
    double* Linv;
    long ldi;
    int* flag;
};
hipError_t chol_diag_batch(const CholJob* jobs, int n, hipStream_t s);
struct DotJob {  // out[i] = V_i . w, i < k
    const double* V;
    long ldv;
    int k;
    const double* w;
    long n;
    double* out;
};
// part: bmdot_part_size(n, kmax, nmax) doubles; every job has k <= kmax, n <= nmax
hipError_t bmdot(const DotJob* jobs, int n, int kmax, long nmax, double* part, hipStream_t s);
long bmdot_part_size(int n, int kmax, long nmax);
struct GemvJob {  // w = beta w + cscale sum_i coef[i] V_i
    const double* V;
    long ldv;
    int k;
    const double* coef;
    double cscale, beta;
    double* w;
    long n;
};
hipError_t bgemv(const GemvJob* jobs, int n, long nmax, hipStream_t s);
struct ThreeJob {  // w -= alpha[0] v + beta u
    double* w;
    const double* v;
    const double* u;
    const double* alpha;
    double beta;
    long n;
};
hipError_t bthree(const ThreeJob* jobs, int n, long nmax, hipStream_t s);
struct ScaleJob {  // y = a x; with a_sq set, a = 1 / sqrt(*a_sq) read on the device
    double a;
    const double* x;
    double* y;
    long n;
    const double* a_sq;
};
hipError_t bscale(const ScaleJob* jobs, int n, long nmax, hipStream_t s);
struct ScaleConvJob {  // y = a x and y32 = (float) y; with a_sq set, a = 1 / sqrt(*a_sq)
    double a;
    const double* x;
    double* y;
    float* y32;
    long n;
    const double* a_sq;
};
hipError_t bscale_conv(const ScaleConvJob* jobs, int n, long nmax, hipStream_t s);
struct ConvJob {  // y = (float) x
    const double* x;
    float* y;
    long n;
};
hipError_t bconvert(const ConvJob* jobs, int n, long nmax, hipStream_t s);
struct MixJob {  // FP32 in; with t.esel/t.osel set, the identity-channel form of wmix_id
    const void* in;
    long isx;
    void* out;
    long osx;
    int X, Bn;
    MixTable t;
    const float* in2;  // identity form only: the FP32 input vector, row stride isx2
    long isx2;
    double* out2;      // identity form only: y, row stride osx2
    long osx2;
};
// Fused Lanczos step for many problems (n even for every job), the batched
// form of lz_step_dots / lz_update_norm.  part: bmdot_part_size(n, kmax, nmax)
// doubles for bstep_dots, n * ceil(nmax / 2048) for bupdate_norm.
struct StepJob {  // w -= alpha[0] V_{k-1} + beta V_{k-2} (k >= 2), then out[i] = V_i . w, i < k;
    double* w;      // with beta_sq set, beta = sqrt(*beta_sq) read on the device
    const double* V;
    long ldv;
    int k;
    const double* alpha;
    double beta;
    long n;
    double* out;
    const double* beta_sq;
};
hipError_t bstep_dots(const StepJob* jobs, int n, int kmax, long nmax, double* part, hipStream_t s);
struct UpdJob {  // w -= sum_{i<k} coef[i] V_i, then out[0] = |w|^2
    double* w;
    const double* V;
    long ldv;
    int k;
    const double* coef;
    long n;
    double* out;
};
hipError_t bupdate_norm(const UpdJob* jobs, int n, long nmax, double* part, hipStream_t s);
struct EyeJob {  // out[0] = max |G - I| over the leading n x n block (row stride ld)
    const double* G;
    long ld;
    int n;
    double* out;
};
// part: 64 * n doubles
hipError_t beye_deviation(const EyeJob* jobs, int n, double* part, hipStream_t s);
struct Copy2Job {  // copy2 of one problem (strided 2-D copy with zero fill)
    const double* in;
    long si0, si1;
    int r, c;
    double* out;
    long so0, so1;
    int ro, co;
};
hipError_t bcopy2(const Copy2Job* jobs, int n, int romax, int comax, hipStream_t s);
struct DiagJob {  // G[i][i] += v, i0 <= i < i1
    double* G;
    long ld;
    int i0, i1;
    double v;
};
hipError_t badd_diag(const DiagJob* jobs, int n, int lenmax, hipStream_t s);
hipError_t bwmix(int tout, const MixJob* jobs, int n, int Xmax, int Bnmax, hipStream_t s);
hipError_t add_diag(double* G, long ldg, int i0, int i1, double v, hipStream_t s);
hipError_t eye_deviation(const double* G, long ldg, int n, double* out, double* part, hipStream_t s);

}  // namespace tn
