// tn_engine.h -- device-side building blocks behind the C ABI (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <array>
#include <map>
#include <string>
#include <vector>

#include "../../../include/tn_abi.h"
#include "tn_kernels.h"

namespace tn {

inline int pad32(int n) { return n <= 0 ? 32 : ((n + 31) / 32) * 32; }

struct TnError {
    int code;
    std::string msg;
};

#define TN_HIP(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            throw ::tn::TnError{TN_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e) + \
                                " at " + __FILE__ + ":" + std::to_string(__LINE__)};   \
    } while (0)
#define TN_REQUIRE(cond, code, msg)                                                    \
    do {                                                                               \
        if (!(cond)) throw ::tn::TnError{(code), std::string(msg)};                    \
    } while (0)

// grow-only device buffer
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n);
    void release();
    template <typename T> T* as() const { return (T*)p; }
};

// host-mapped coherent memory for the per-iteration readback of a few
// scalars: the device publishes values and a sequence number, the host polls
struct Scalars {
    double* h = nullptr;              // host view: TN_PUBLISH_MAX doubles, then the flag
    double* d = nullptr;              // device view of the same memory
    unsigned long long* flag_h = nullptr;
    unsigned long long* flag_d = nullptr;
    unsigned long long seq = 0;
};

// Sparse MPO mixing table on the device (see wmix in tn_kernels.h)
struct DevMix {
    DBuf buf;
    MixTable t{};
    // key: the W values themselves (a caller may refill one host buffer with
    // another operator of the same shape, so its address identifies nothing)
    std::vector<double> wkey;
    const double* wsrc = nullptr;
    int op = -1, D1p = -1, D2p = -1, p = -1, dw1 = -1, dw2 = -1, wl = -1, wr = -1;
};
// MIX_HEFF_T: the Heff mixing of heff_apply, written transposed for an NT
// stage 3 (T2t[(m,b)][(w,a')]), with esel/osel always present (identity
// channels optional); MIX_HEFF: the [w][a'][m][b] form (batched path).
enum MixOp { MIX_HEFF = 0, MIX_ENVL = 1, MIX_ENVR = 2, MIX_HEFF_T = 3 };
// W host array [dw1][dw2][p][p]; builds (or reuses) the table for the op/dims.
// MIX_HEFF with wl >= 0 or wr >= 0: table for the operator whose left channel
// wl / right channel wr is the identity (first or last channel), with the
// stage-1 output holding only the other right channels and the stage-3 input
// only the other left channels (see HeffOp).
void mix_build(DevMix& m, int op, const double* W, int dw1, int dw2, int p, int D1p, int D2p,
               hipStream_t s, int wl = -1, int wr = -1);

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
};

}  // namespace tn

struct tn_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // scratch
    tn::DBuf T1, T2, xlo, part, krylov, lanczos_small, svd_X, svd_V, svd_U, svd_Vt, svd_s,
        svd_perm, svd_pairs, svd_flag, svd_A, spare, pack, qr_W, qr_G, qr_L, qr_T, qr_linv, qr_Q;
    // preconditioned SVD: R, Q_h, reflectors (then Q_h J), T factors + GEMM work space
    tn::DBuf svd_R, svd_Q, svd_H, svd_W, svd_X2;
    tn::Scalars sc;
    hipEvent_t ev[8] = {};
    // GEMM autotuning (off unless requested): best (tile code, k-split) per
    // problem shape, measured on first use
    bool autotune = false;
    std::map<std::array<long, 9>, std::pair<int, int>> tune;
    tn::DBuf tune_C, splitk;
    // Davidson workspace and the preconditioner network (params + norm, FP32)
    tn::DBuf dav_W, dav_vec, dav_wd, mlp;
    // identity-channel check of the environments (4 device doubles)
    tn::DBuf ident;
    // deflation of the lowest eigenvector for the second eigenpair (dot + partials)
    tn::DBuf defl;
    // device address of this device's fixed work space (allocation-free apply)
    void* noalloc = nullptr;
    // stats whose Heff event timings are recorded but not yet read
    tn_sweep_stats* heff_pending = nullptr;
    // nonzero: the next heff_apply first scales its input in place by this
    // factor (within its first pass over it) and resets it; set by the Lanczos
    // driver right before the apply of a new basis vector
    double x_prescale = 0.0;
    // seam state (tn_site_bind)
    int b_d = 0, b_D1 = 0, b_D2 = 0, b_Dw1 = 0, b_Dw2 = 0, b_prec = 0;
    int b_wl = -1, b_wr = -1;  // identity channels of the bound environments
    std::vector<double> b_W;
    tn::DBuf b_L, b_R, b_Llo, b_Rlo, b_x, b_y, b_M;
    tn::DevMix b_mix_heff, b_mix_envl, b_mix_envr;
    bool bound = false;
};

namespace tn {

// GEMM on the context's stream.  With g.tile == 0 and ctx->autotune set, the
// tile and k-split come from the per-shape table (tuned on first use by timing
// every candidate into scratch memory); otherwise the built-in choice.
void gemm_run(tn_ctx* c, const GemmDesc& g);
// the same GEMM with an explicit k-split (ks > 1 needs batch == 1, K % (64 ks) == 0)
void gemm_split(tn_ctx* c, const GemmDesc& g, int ks);
// (tile code, k-split) gemm_run would use for g (tuning it first when
// autotuning is on and the shape is new)
std::pair<int, int> gemm_choice(tn_ctx* c, const GemmDesc& g);

// --- effective Hamiltonian operator (internal padded layout) -------------
// x, y: [D1p][p][D2p]   L: [D1p][Dw1][D1p]   R: [Dw2][D2p][D2p]
struct HeffOp {
    int p, D1p, D2p, Dw1, Dw2;
    const double* L;
    const double* R;
    const void* Llo;  // low precision copies (nullptr when precision == FP64)
    const void* Rlo;
    long Lps = 0, Rps = 0;  // TN_BF16X3 copies: element stride between the three planes
    int precision;
    const DevMix* mix;  // MIX_HEFF table for (D1p, D2p, wl_id, wr_id)
    int tile;
    // identity environment channels (-1: none).  L[a][wl_id][a'] = delta(a,a')
    // and R[wr_id][b][b'] = delta(b,b') on the logical block (zero padding):
    // stage 1 skips channel wr_id (its slice is x itself), stage 3 skips
    // channel wl_id (its contribution is the mixed slice copied into y).
    // Each must be 0 or the last channel of its side, with Dw >= 2.
    int wl_id = -1, wr_id = -1;
    long n() const { return (long)D1p * p * D2p; }
    double flops_alg(int D1, int D2) const;  // reference formula on logical dims
    // flops of the two GEMM stages as launched (identity channels excluded)
    double flops_gemm(int D1, int D2) const;
};
// Work space of one Heff apply supplied by the caller.  With it the apply
// allocates nothing: the stage buffers come from here and every GEMM runs in
// one launch with the tile of the tuning table (or the built-in choice),
// never a k-split and never a tuning run.
struct HeffScratch {
    void* T1;
    size_t T1_bytes;
    void* T2;
    size_t T2_bytes;
    void* xlo;  // reduced-precision copy of x (unused in FP64)
    size_t xlo_bytes;
};
// bytes of each part (256-byte multiples) and their sum
size_t heff_scratch_bytes(const HeffOp& op, size_t* t1, size_t* t2, size_t* xlo);
// GEMM with the table's tile (or the built-in one), one launch, no tuning
void gemm_run_fixed(tn_ctx* c, const GemmDesc& g);
// y = Heff x (fp64 in/out).  timed: HIP events around the stages (read by
// heff_timing_read / heff_stage_times).  scr: caller work space (see HeffScratch),
// nullptr = the context's buffers.
void heff_apply(tn_ctx* c, const HeffOp& op, const double* x, double* y, bool timed,
                const HeffScratch* scr = nullptr);
void heff_timing_read(tn_ctx* c, double* gemm_ms, double* total_ms);
// per-stage times of the last timed apply (ms): s1, mix, s3
void heff_stage_times(tn_ctx* c, double* s1, double* mix, double* s3);

// --- environment updates ---------------------------------------------------
// left : Lout[D2p][Dw2][D2p]  from Lin[D1p][Dw1][D1p] and A[D1p][p][D2p]
// right: Rout[Dw1][D1p][D1p]  from Rin[Dw2][D2p][D2p] and A
void env_left(tn_ctx* c, int prec, int p, int D1p, int D2p, int Dw1, int Dw2, const DevMix& mix,
              const double* Lin, const void* Lin_lo, const double* A, double* Lout, int tile);
void env_right(tn_ctx* c, int prec, int p, int D1p, int D2p, int Dw1, int Dw2, const DevMix& mix,
               const double* Rin, const void* Rin_lo, const double* A, double* Rout, int tile);
// store a low precision copy of an fp64 buffer (TN_BF16X3: three bfloat16
// planes ps elements apart, ps = 0 meaning n)
void make_lowprec(tn_ctx* c, int prec, const double* src, void* dst, long n, long ps = 0);
// bytes per element of a low precision copy (all planes)
inline size_t lowprec_bytes(int prec) { return prec == TN_FP32 ? 4 : (prec == TN_BF16X3 ? 6 : 2); }

// --- Lanczos ---------------------------------------------------------------
struct LanczosResult {
    double e0 = 0, e1 = 0, residual = 0;
    int n_matvec = 0;
    int n_full = 0;  // iterations with a Gram-Schmidt pass against the whole basis
    bool converged = false;
};
// lowest two eigenvalues of tridiag(b, a, b) (k x k) and the eigenvector z of the lowest
// (t1 only when want_t1; else it is set to t0)
void tridiag_lowest(int k, const double* a, const double* b, double* t0, double* t1, double* z, bool want_t1 = true);
// generic operator callback: y = A x (device fp64)
typedef void (*ApplyFn)(tn_ctx* c, const void* op, const double* x, double* y);
// converged when |H x - theta x| < tol or < tol_rel*|theta| (tol_rel = 0: off)
// partial = true: local re-orthogonalisation against the last two Krylov
// vectors every iteration, and a pass against the whole basis only when the
// estimated loss of orthogonality exceeds sqrt(eps_op), eps_op the relative
// rounding of the operator (Simon's omega recurrence); false: the pass
// against the whole basis every iteration
//
// prescale: fn applies the Heff through heff_apply, which honours
// tn_ctx::x_prescale; the scaling v_{j+1} = w / beta_j is then left to the
// next apply's first pass over its input instead of a separate kernel
LanczosResult lanczos(tn_ctx* c, ApplyFn fn, const void* op, long N, double* x, double tol,
                      int max_krylov, int max_restarts, double tol_rel = 0.0, bool partial = false,
                      double eps_op = 0x1p-52, bool prescale = false);

// --- Davidson with a diagonal / learned preconditioner (tn_precond.hip) ---
struct DavidsonResult {
    double e0 = 0, e1 = 0, residual = 0;
    int n_matvec = 0, iters = 0;
    bool converged = false;
};
// geo: the operator's shapes and FP64 environments (for the Heff diagonal),
// D1 x D2 the logical bond dimensions;
// Wd: device [Dw1][Dw2][p] diagonal of the W blocks; kind 0 none, 1 diagonal,
// 2 network (ctx->mlp), 3 network on the first correction and diagonal after;
// converged when |H u - theta u| < tol.
DavidsonResult davidson(tn_ctx* c, ApplyFn fn, const void* op, const HeffOp& geo, int D1, int D2, const double* Wd,
                        double* x, double tol, int mmax, int max_iter, int kind, double delta);
void heff_diag(tn_ctx* c, const HeffOp& op, const double* Wd, double* diag);
void bond_weights(tn_ctx* c, const double* u, int p, int D1p, int D2p, double* laml, double* lamr);
void sym_eig_small(int m, const std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z);

// --- SVD -----------------------------------------------------------------
// Thin SVD of a device matrix view A(i,j) = A[i*rs + j*cs], r x c.
// Writes U (row-major r x k, ldu), Vt (row-major k x c, ldvt), s (host, k).
// U/Vt buffers must hold ru x ldu / ku x ldvt and are zero filled first over
// [0, zr) x [0, ldu) and [0, zk) x [0, ldvt).
int svd_device(tn_ctx* c, const double* A, long rs, long cs, int r, int c_, double* U, long ldu,
               int zr, double* Vt, long ldvt, int zk, std::vector<double>& s);

// Shifted CholeskyQR (two passes, three when needed) of a row-major m x n
// matrix A (lda), m >= n, both multiples of 32, leading nl columns logical
// (the rest zero padding).  On success Q (row-major m x n) has orthonormal
// logical columns, zero padded columns, and Lacc (n x n row-major, lower)
// satisfies A = Q Lacc^T.  Returns false if max|Q^T Q - I| > tol after the
// passes (the caller falls back to SVD).  *shifted = 1 when the first pass
// needed the shift.  With triangular = false a first pass that is already
// nearly orthogonal is finished by a polar correction (Lacc then full).
bool cholqr_device(tn_ctx* c, const double* A, long lda, int m, int n, int nl, double* Q,
                   double* Lacc, double tol, int* shifted = nullptr, bool triangular = true);

void sync(tn_ctx* c);
// stream-ordered copies that return when the data has arrived
void h2d(tn_ctx* c, void* dst, const void* src, size_t bytes);
void d2h(tn_ctx* c, void* dst, const void* src, size_t bytes);
// n (<= TN_PUBLISH_MAX) doubles device -> host through the published scalars
// (no stream synchronisation: the host polls the completion flag)
void d2h_small(tn_ctx* c, double* dst, const double* src, int n);

}  // namespace tn
