/*
 * tn_abi.h -- C ABI of the MI355X (gfx950) DMRG hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * an int status: 0 = ok, < 0 = error (HIP failure, bad argument, bad state),
 * > 0 = numerical warning (e.g. eigensolver not converged).  tn_last_error()
 * gives the message of the last failure on a context.  One host thread per
 * context; host buffers stay owned by the caller; calls are synchronous.
 *
 * Layouts at this boundary are the reference's (0-based):
 *   site tensor   psi[m + d*(i1 + D1*i2)]         (Containers/Matrix6.H:110-111)
 *   environment   F[w + Dw*(i + D*j)], i bra, j ket (Containers/Vector3.H:99-102)
 *   MPO block     W[((w1*Dw2 + w2)*d + m)*d + n], m bra, n ket
 *                                                  (Operators/OperatorValuedMatrix.H:11)
 *   dense matrix  column-major (LAPACK convention of the reference solvers)
 */
#ifndef TN_ABI_H
#define TN_ABI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tn_ctx tn_ctx;
typedef struct tn_chain tn_chain;

enum tn_status {
    TN_OK = 0,
    TN_WARN_NOT_CONVERGED = 1,
    TN_ERR_HIP = -1,
    TN_ERR_ARG = -2,
    TN_ERR_STATE = -3,
    TN_ERR_NOMEM = -4
};

/* arithmetic of the contractions (Heff apply / environment updates).
 * TN_BF16X3: every FP64 operand is split into three bfloat16 pieces
 * (x = x0 + x1 + x2, about 24 significant bits) and the contraction sums the
 * six leading cross products x_i y_j (i + j <= 2) on the bfloat16 matrix
 * cores with FP32 accumulation: FP32-class accuracy at bfloat16 throughput.
 * Heff only (tn_sweep_opts.precision, tn_site_bind); environment updates
 * take FP64, FP32 or BF16. */
enum tn_precision { TN_FP64 = 0, TN_FP32 = 1, TN_BF16 = 2, TN_BF16X3 = 3 };

/* ---- context ------------------------------------------------------------ */
/* Replaces nothing in the reference (it has no device); owns the HIP stream
 * and scratch memory.  One context per chain or per batch, no globals. */
int tn_ctx_create(int device, tn_ctx** ctx);
int tn_ctx_destroy(tn_ctx* ctx);
/* GEMM autotuning for this context (no reference counterpart: the reference's
 * contractions are fixed loops).  When on, every matrix-core GEMM whose shape
 * is seen for the first time is timed over the tile shapes and k-splits of the
 * kernel menu (into scratch memory) and the fastest is kept for that shape.
 * Off by default: results then do not depend on timing. */
int tn_ctx_set_autotune(tn_ctx* ctx, int on);
/* Text table of the tuned shapes ("f32->f64 KT M=.. N=.. K=.. : tile 128x64
 * ksplit 2" per line) into buf (len bytes); returns the number of shapes. */
int tn_ctx_tune_report(tn_ctx* ctx, char* buf, int len);
/* Tuning table persistence: one line per shape, "tin tout akm bkm M N K batch
 * accumulate tile ksplit" (integers).  load merges into the context's table
 * (returns the number of entries read, < 0 on error); save writes it. */
int tn_ctx_tune_load(tn_ctx* ctx, const char* path);
int tn_ctx_tune_save(tn_ctx* ctx, const char* path);
const char* tn_last_error(const tn_ctx* ctx);
int tn_version(void);
/* number of visible HIP devices (0 when none or when the runtime fails) */
int tn_device_count(void);

/* ---- reference solver seam (host pointers, reference layouts) ---------- */
/* Binds the operands of one site's effective Hamiltonian: W (Dw1 x Dw2 blocks
 * of d x d), left environment L (Dw1, D1, D1) and right environment R
 * (Dw2, D2, D2).  Replaces the inputs of MPSSite::GetHeff
 * (TensorNetworksImp/MPS/MPSSite_Contractions.C:155-193). */
int tn_site_bind(tn_ctx* ctx, int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                 const double* L, const double* R, int precision);
/* y = Heff x, N = d*D1*D2: the matrix-free product behind
 * SparseEigenSolverClient<T>::DoMatVecContraction(int N, const T* x, T* y)
 * (NumericalMethods/SparseEigenSolver.H:20-21), replacing the dense Heff of
 * MPSSite::GetHeff. */
int tn_heff_apply(tn_ctx* ctx, const double* x, double* y, int64_t N);
/* Environment update through the bound site operator: dir 0 = IterateLeft_F
 * (uses the bound L, writes F of shape (Dw2, D2, D2)), dir 1 = IterateRightF
 * (uses the bound R, writes (Dw1, D1, D1)).  M is the site tensor.
 * TensorNetworksImp/MPS/MPSSite_Contractions.C:292-400. */
int tn_env_update(tn_ctx* ctx, int dir, const double* M, double* Fout);
/* The same product with x and y in device memory of the context's device
 * (reference layout, N doubles each).  The data must be complete before the
 * call (the library's stream does not wait on other streams); returns when y
 * is written.  No PCIe traffic: the layout change runs on the device. */
int tn_heff_apply_device(tn_ctx* ctx, const double* x_dev, double* y_dev, int64_t N);
/* tn_heff_apply_device that allocates nothing (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc.  Every
 * intermediate (x and y in the internal layout, the stage buffers) lives in a
 * fixed 64 MiB work space per GPU, a device array the library holds from load
 * until the process ends; its GEMMs run in one launch each with the tuning
 * table's tile (no k-split, never a tuning run).  A site whose work space
 * exceeds 64 MiB (for FP64 about d*D1*D2*(2 + 2*Dw) doubles, e.g. d=3 D=400
 * Dw=5) is rejected with TN_ERR_ARG.  The site must be bound with
 * tn_site_bind first (binding allocates).  Calls that overlap on one GPU
 * (several contexts or threads at once) share the work space and race. */
int tn_heff_apply_device_noalloc(tn_ctx* ctx, const double* x_dev, double* y_dev, int64_t N);
/* Lowest eigenpair(s) of the bound Heff by Lanczos with full FP64
 * re-orthogonalisation; replaces LapackEigenSolver<dcmplx>::Solve(Heff, eps, 2)
 * (NumericalMethods/LapackEigenSolver.C:62-109: IL=1, IU=2).  evec holds nev
 * columns of N doubles (column k at evec + k*N): on input column 0 is the
 * start vector (all zero or non-finite -> seeded random), on output column k
 * is the eigenvector of evals[k] (unit norm), evals ascending, nev <= 2.  The
 * second pair comes from Lanczos on H + sigma x0 x0^T (x0 the first vector),
 * sigma above the spread of the first Ritz values.  n_matvec counts both
 * solves.  Returns TN_WARN_NOT_CONVERGED if maxit is reached. */
int tn_eig_lowest(tn_ctx* ctx, int64_t N, int nev, double tol, int maxit, double* evals,
                  double* evec, int* n_matvec);
/* Thin SVD A = U diag(s) Vt of a column-major m x n matrix, k = min(m,n):
 * U (m x k), s (k, descending), Vt (k x n), column-major; U and Vt have
 * orthonormal columns/rows even for zero singular values.  FP64 one-sided
 * block Jacobi on the device; for k >= 128 with columns far from orthogonal
 * (largest cosine >= 0.1) the Jacobi runs on the triangular factor of a
 * Householder QR of the norm-sorted columns (the environment variable
 * TN_SVD_PLAIN=1 turns that off).  Replaces LapackSVDSolver<T>::Solve
 * (NumericalMethods/LapackSVDSolver.C:42-90). */
int tn_svd(tn_ctx* ctx, int m, int n, const double* A, double* U, double* s, double* Vt);
/* Thin QR / LQ of a column-major m x n matrix on the device (shifted
 * CholeskyQR, FP64); replaces LapackQRSolver<T>::SolveThinQR / SolveThinLQ as
 * MPSSite::NormalizeQR calls them (TensorNetworksImp/MPS/MPSSite_Normalization.C:80-107).
 * mode 0: A = Q R with m >= n, Q (m x n) orthonormal columns, R (n x n) upper
 * triangular with a non-negative diagonal; mode 1: A = L Q with m <= n,
 * L (m x m) lower triangular, Q (m x n) orthonormal rows.  Column-major
 * outputs.  Returns TN_WARN_NOT_CONVERGED when max|Q Q^T - I| (the
 * orthogonality of the computed factor) stays above 1e-11, < 0 on error
 * (including m < n for mode 0 and m > n for mode 1, the forms LAPACK returns
 * with a rectangular triangular factor). */
int tn_qr(tn_ctx* ctx, int mode, int m, int n, const double* A, double* Q, double* R);
/* Truncation rule of SVCompressorImp<T>::Compress
 * (TensorNetworksImp/SVCompressorImp.C:19-73): returns the kept dimension,
 * the discarded weight sum(s^2) and the squared-norm ratio rescale factor. */
int tn_svcompress_rank(int N, const double* s, int Dmax, double eps, double* discarded,
                       double* rescale);

/* ---- MPO canonical form and compression (host) ---------------------------
 * MPO::CanonicalForm() and MPO::Compress(Parker, Dmax, eps)
 * (TensorNetworks/MPO.C:93-128; Operators/SiteOperatorImpCompress.C:16-100 ->
 * MatrixO::BlockQX / BlockSVD, Operators/OperatorValuedMatrix.C:437-537,786-840),
 * as Hamiltonian::CreateH2Operator applies them (TensorNetworks/Hamiltonian.C:8-16).
 * The MPO is given like tn_chain_create's (dw1, dw2, W concatenated) in
 * regular form (form 0 = RegularUpper, 1 = RegularLower: the identity
 * channels first/last on every bond).  ops: 1 = canonical form (block QR
 * sweeps left then right), 2 = compression (block QR + SVD of the interior
 * block, left then right, singular values cut by the SVCompressor rule
 * (Dmax, eps); Dmax = 0 no cap), 3 = both.  Writes the new bond dims into
 * dw1_out/dw2_out (L each) and the blocks into w_out when cap (doubles) is
 * large enough; returns the number of doubles the result needs (>= 0), or
 * < 0 on error (also for an MPO not in regular form).  *terror (optional):
 * sqrt(sum of discarded s^2) / (L-1), as MPO::Compress reports it.  Host
 * only: needs no device. */
int tn_mpo_reduce(int L, int d, const int* dw1, const int* dw2, const double* wcat, int form, int ops, int Dmax,
                  double eps, int* dw1_out, int* dw2_out, double* w_out, int64_t cap, double* terror);

/* ---- device-resident chain (the sweep) ---------------------------------- */
/* Per-site MPO: dw1[i], dw2[i] and W blocks concatenated in site order.
 * Sites start as in MPSImp::InitSitesAndBonds (TensorNetworksImp/MPS/MPSImp.C:69-90):
 * site 0 is 1 x D, bulk D x D, site L-1 is D x 1; bonds hold 1/D.
 * Device memory is sized for bond dimensions up to Dcap; growing beyond it
 * (tn_chain_increase_bond_dims) reallocates and invalidates the environment
 * caches, which tn_chain_load_caches rebuilds. */
int tn_chain_create(tn_ctx* ctx, int L, int d, const int* dw1, const int* dw2,
                    const double* wcat, int D, int Dcap, tn_chain** chain);
/* Destroy a chain; call before tn_ctx_destroy of its context. */
int tn_chain_destroy(tn_chain* chain);
/* MPSSite::InitializeWith (TensorNetworksImp/MPS/MPSSite.C:84-140); state
 * 0 = Product, 1 = Random (seeded U(-1,1) / sqrt(d D1 D2)), 2 = Neel. */
int tn_chain_initialize(tn_chain* chain, int state, uint64_t seed);
int tn_chain_get_dims(const tn_chain* chain, int* D1, int* D2);
int tn_chain_get_site(tn_chain* chain, int site, double* M); /* reference layout */
int tn_chain_set_site(tn_chain* chain, int site, int D1, int D2, const double* M);

/* Options of the local update.  precision: Heff arithmetic of the Lanczos
 * iterations; env_precision: environment-update arithmetic; refine: restart
 * the Lanczos from the reduced-precision Ritz vector with the
 * refine_precision operator until refine_tol (0/1). */
typedef struct tn_sweep_opts {
    double eig_tol;      /* Lanczos residual tolerance |H x - theta x| */
    int max_krylov;      /* Krylov dimension cap (restarted beyond) */
    int max_restarts;
    int precision;       /* enum tn_precision */
    int env_precision;   /* enum tn_precision */
    int refine;          /* FP64 refinement after reduced-precision Lanczos */
    double refine_tol;
    int gemm_tile;       /* 0 = automatic tile choice */
    int gauge;           /* 0 = SVD gauge moves (MPSSite::SVDNormalize, bond singular
                            values kept); 1 = QR gauge moves (MPSSite::NormalizeQR,
                            MPSSite_Normalization.C:80-107: bond values cleared as in
                            Bond::ClearSingularValues), by CholeskyQR2 with fallback
                            to the SVD when it is not accurate */
    int refine_precision;  /* operator precision of the refinement (FP64 or FP32),
                              used when refine=1 and it is finer than precision */
    int eig_method;      /* 0 = Lanczos; 1 = Davidson (basis max_krylov, at most
                            max_krylov*(1+max_restarts) iterations), an alternative to
                            the reference's SparseEigenSolver back ends */
    int precond;         /* Davidson correction: 0 none, 1 diagonal (Heff diagonal),
                            2 learned (network set by tn_ctx_set_mlp) at every step,
                            3 learned on the first correction from the site tensor and
                            diagonal after; with 2 and 3 Davidson gets max_krylov
                            iterations, then plain Lanczos continues from its Ritz
                            vector if it has not converged */
    double precond_delta;  /* safeguard of the diagonal shift |diag - theta| >= delta */
    int skip_identity;   /* 1 = an environment channel equal to the identity (the
                            "nothing placed yet" / "all placed" MPO channel of an
                            orthonormal block) is applied as a copy instead of a GEMM
                            slice; checked per local solve against the identity
                            (1e-12 for FP64 environments, 1e-4 for FP32 ones) */
    int time_heff;       /* 1 = HIP events around the stages of every Heff apply
                            (tn_sweep_stats.heff_ms / heff_gemm_ms); 0 = counters only,
                            no events in the stream */
    int reorth;          /* Lanczos re-orthogonalisation: 0 = against the whole Krylov
                            basis every iteration; 1 = partial (Simon's omega recurrence):
                            against the last two vectors every iteration and against the
                            whole basis only when the estimated loss of orthogonality
                            exceeds sqrt(u), u the unit roundoff of the Heff precision */
} tn_sweep_opts;

typedef struct tn_sweep_stats {
    int64_t n_matvec;          /* Heff applications (all precisions) */
    int64_t n_matvec_fp64;     /* of which in FP64 */
    int64_t n_env;             /* environment updates */
    int64_t n_svd;             /* bond SVDs */
    double heff_flops;         /* algorithmic flops of all Heff applies */
    double heff_ms;            /* device time of all Heff applies (HIP events) */
    double heff_gemm_ms;       /* device time of the two GEMM stages only */
    double heff_gemm_flops;    /* flops of the two GEMM stages as launched (identity
                                  environment channels excluded) */
    double env_flops, env_ms;
    double svd_ms, lanczos_ms, wall_ms;
    double energy;             /* last site eigenvalue */
    int64_t n_qr;              /* QR gauge moves */
    int64_t n_qr_fallback;     /* QR gauge moves that fell back to the SVD */
    int64_t n_qr_shifted;      /* QR gauge moves whose first pass needed the shift */
    int64_t n_ident_left;      /* local solves whose left environment had an identity channel */
    int64_t n_ident_right;     /* ... right environment */
    int64_t n_precond_fallback;  /* learned-correction Davidson solves finished by Lanczos */
    int64_t n_full_reorth;     /* Lanczos iterations with a Gram-Schmidt pass against the whole
                                  Krylov basis (every iteration with reorth = 0) */
    int64_t n_solve_converged; /* local eigensolves that met eig_tol (refine_tol when refined) */
    int64_t n_solve_capped;    /* local eigensolves stopped by the iteration budget
                                  max_krylov * (1 + max_restarts) above the tolerance */
} tn_sweep_stats;

void tn_sweep_opts_default(tn_sweep_opts* o);
/* MPS::Normalize(Direction) via SVD gauge moves, no compression
 * (TensorNetworks/MPS.C:8, TensorNetworksImp/MPS/MPSSite_Normalization.C:45-78).
 * dir 0 = DLeft, 1 = DRight. */
int tn_chain_normalize(tn_chain* chain, int dir);
/* MPS::NormalizeQR(Direction) (TensorNetworksImp/MPS/MPSSite_Normalization.C:80-107):
 * QR gauge moves (CholeskyQR2 on the device, SVD fallback); bond values cleared. */
int tn_chain_normalize_qr(tn_chain* chain, int dir);
/* MPSImp::LoadHeffCaches (TensorNetworksImp/MPS/MPSImp_Variational.C:120-124) */
int tn_chain_load_caches(tn_chain* chain, const tn_sweep_opts* opts);
/* MPSImp::Sweep (TensorNetworksImp/MPS/MPSImp_Variational.C:64-76): Refine every
 * site in direction dir (0 = sites 0..L-1, 1 = L-1..0). */
int tn_chain_sweep(tn_chain* chain, int dir, const tn_sweep_opts* opts, tn_sweep_stats* st);
/* MPSImp::Refine (MPSImp_Variational.C:79-99) for one site. */
int tn_chain_refine_site(tn_chain* chain, int site, int dir, const tn_sweep_opts* opts,
                         tn_sweep_stats* st);
/* MPSImp::IncreaseBondDimensions (TensorNetworksImp/MPS/MPSImp.C:123-147) */
int tn_chain_increase_bond_dims(tn_chain* chain, int D);
/* MPSImp::GetMaxDeltaE / GetMaxDeltal (MPSImp_Variational.C:152-173) */
int tn_chain_deltas(const tn_chain* chain, double* max_de, double* max_dl);
/* Bond b (between sites b and b+1): normalised singular values (n <= cap). */
int tn_chain_bond_sv(const tn_chain* chain, int b, double* sv, int cap);
/* MPSImp::GetExpectation(const MPO*) (TensorNetworksImp/MPS/MPSImp_ExpectationsAndDMs.C:33-45)
 * for any MPO given like tn_chain_create's; FP64 left zipper on the device. */
int tn_chain_expectation(tn_chain* chain, const int* dw1, const int* dw2, const double* wcat,
                         double* value);
/* MPS::GetNormStatus() (TensorNetworksImp/MPS/MPSSite.C:173-211): per site
 * 'A' left-normalised (sum_{m,i1} A^T A = 1), 'B' right-normalised, 'I' both,
 * 'M' neither, within eps (elementwise); out receives L characters and a NUL. */
int tn_chain_norm_status(tn_chain* chain, double eps, char* out);
/* Checkpoint of a chain (no reference counterpart; the reference's MPS::Clone,
 * TensorNetworksImp/MPS/MPSImp.C:38-53, copies in memory only): MPO, site
 * tensors (reference layout), bond values and per-site energies, with
 * flags & 1 also the FP64 environment caches, and `meta_bytes` bytes of the
 * caller's own data (e.g. the position in a schedule).  Written to path.tmp,
 * then renamed; a hash guards the content. */
int tn_chain_save(tn_chain* chain, const char* path, const void* meta, int64_t meta_bytes, int flags);
/* Recreate a chain from a checkpoint on ctx (bond capacity max(Dcap, the
 * saved one)).  Without saved environments call tn_chain_load_caches before
 * sweeping.  meta receives the first min(meta_cap, length) bytes of the
 * caller's data and *meta_bytes its full length. */
int tn_chain_load(tn_ctx* ctx, const char* path, int Dcap, tn_chain** chain, void* meta, int64_t meta_cap,
                  int64_t* meta_bytes);
/* Two-site update of bond (site, site+1) with the SVCompressor rule (Dmax, eps):
 * an extension (the reference sweep is one-site only). */
int tn_chain_sweep2(tn_chain* chain, int dir, int Dmax, double eps_sv,
                    const tn_sweep_opts* opts, tn_sweep_stats* st, double* max_trunc);

/* ---- diagnostics (no counterpart in the reference) ---------------------- */
/* Timing of `reps` Heff applications at one site of the chain in the given
 * precision (HIP events on the library stream): average device time of the
 * first GEMM stage, the MPO stage and the second GEMM stage, in ms, and the
 * flops of one launch of each GEMM stage as launched.  skip_identity = 1 uses
 * the operator a sweep with tn_sweep_opts.skip_identity = 1 would build
 * (identity environment channels found with env_precision's tolerance).
 * The site tensor is left unchanged. */
int tn_chain_heff_bench(tn_chain* chain, int site, int precision, int reps, int skip_identity, int env_precision,
                        double* s1_ms, double* mix_ms, double* s3_ms, double* s1_flops, double* s3_flops);

/* One batched GEMM through the device kernel, for numerics tests of the
 * matrix-core path: C[b] = A[b] B[b] with operands rounded to tin
 * (enum tn_precision values 0 fp64, 1 fp32, 2 bf16, 3 split bf16 -- the
 * last with a_kmajor = b_kmajor = 1 only) and the result stored as
 * tout before conversion back to double.  a_kmajor: A[b] is M x K row-major
 * (else K x M); b_kmajor: B[b] is N x K row-major (else K x N).  M, N multiples
 * of 32, K of 32.  tile: 0 automatic, else a forced tile code.  ksplit > 1
 * splits K into that many partial products summed by a reduction kernel
 * (batch 1, K a multiple of 64 ksplit). */
int tn_gemm_test(tn_ctx* ctx, int tin, int tout, int a_kmajor, int b_kmajor, int M, int N, int K,
                 int batch, const double* A, const double* B, double* C, int tile, int ksplit);

/* ---- batched chains (no reference counterpart: the reference sweeps one
 * MPS at a time) ----
 * A batch advances many chains of one context in lockstep: pass dir (0 left
 * to right, 1 right to left) gives round t to site t (or L-1-t) of every chain
 * with L > t, and every stage of a round is one grouped launch over the
 * chains (grouped persistent GEMMs with per-shape tiles, batched Lanczos
 * vector kernels, batched CholeskyQR).  The per-site algorithm is that of
 * tn_chain_sweep with gauge = 1 (QR) and FP32 contractions: opts must have
 * precision = env_precision = TN_FP32 and max_krylov <= the batch's.  The
 * chains stay owned by the caller, need their caches loaded (FP32 or FP64
 * environments) and must outlive the batch.  energies (optional, n doubles)
 * receives each chain's last Ritz value. */
typedef struct tn_batch tn_batch;
int tn_batch_create(tn_ctx* ctx, int n, tn_chain* const* chains, int max_krylov, tn_batch** batch);
int tn_batch_destroy(tn_batch* batch);
int tn_batch_sweep(tn_batch* batch, int dir, const tn_sweep_opts* opts, tn_sweep_stats* stats, double* energies);

/* ---- learned Davidson preconditioner (no reference counterpart) ----
 * Network 4 -> 64 -> 64 -> 1 (tanh, FP32): params TN_MLP_NP floats
 * (W1[64][4], b1[64], W2[64][64], b2[64], w3[64], b3), norm 8 floats
 * (mu[4], inv_sd[4]) standardising the raw features
 * x = [asinh((diag - theta)/delta), (theta - mean(diag))/sd(diag), log(lam_l + 1e-10),
 * log(lam_r + 1e-10)] (diagonal statistics over the logical elements).
 * The Davidson correction is t = g * r / (theta - diag) with g the output
 * clamped to [-16, 16]. */
#define TN_MLP_NP 4545
/* Load the network used by eig_method = 1, precond = 2 on this context. */
int tn_ctx_set_mlp(tn_ctx* ctx, const float* params, const float* norm);
/* g[i] = network(X[i]) for raw features X (n x 4). */
int tn_mlp_forward(tn_ctx* ctx, const float* params, const float* norm, int64_t n, const double* X, double* g);
/* Mean loss (1/n) sum (g_i a_i - t_i)^2 and its gradient (FP32 kernels). */
int tn_mlp_loss_grad(tn_ctx* ctx, const float* params, const float* norm, int64_t n, const double* X,
                     const double* a, const double* t, double* loss, float* grad);
/* Adam training on the device: `steps` minibatches of `batch` samples drawn
 * (seeded) from the n samples; params updated in place.  loss[0] / loss[1]:
 * mean loss over all samples before / after. */
int tn_mlp_train(tn_ctx* ctx, float* params, const float* norm, int64_t n, const double* X, const double* a,
                 const double* t, int steps, int batch, double lr, uint64_t seed, double* loss);
/* Training samples from the local problem of site `site` (current
 * environments, FP64 operator): the exact local ground state u0 is found by
 * Lanczos, then `iters` Davidson iterations with the diagonal correction run
 * from the current site tensor; after each, nsamp elements (seeded) are
 * recorded as raw features X (4), a = r/(theta - diag) and the exact
 * correction t = u0/(u.u0) - u, both divided by |r|.  Arrays hold
 * iters*nsamp entries; *nout = entries written.  The chain is unchanged. */
int tn_chain_precond_samples(tn_chain* ch, int site, int iters, int nsamp, uint64_t seed, double delta,
                             double* X, double* a, double* t, int* nout);

/* Round trip of n (<= 16384) doubles host -> device -> host through the
 * polled read-back the Lanczos iterations use (values published to host-mapped
 * memory, then a completion flag), for a value test of that path. */
int tn_publish_test(tn_ctx* ctx, int n, const double* in, double* out);

/* Plain device copy of n doubles (read n*8 bytes, write n*8 bytes) by the
 * library's streaming kernel, once, into freshly allocated buffers; *ms = its
 * duration.  A byte-counter run uses it as the known-size calibration. */
int tn_copy_bench(tn_ctx* ctx, int64_t n, double* ms);

/* The QR gauge kernel on its own, for numerics tests: A (row-major m x n,
 * real, m >= n) = Q R with Q (m x n) orthonormal columns and R (n x n) upper
 * triangular, by shifted CholeskyQR on the device (the factorisation behind
 * tn_sweep_opts.gauge = 1; reference MPSSite_Normalization.C:80-107 uses a
 * Householder QR for the same step).  Returns 0 when the factorisation met
 * max|Q^T Q - I| <= tol, 1 when it did not (Q, R still written), < 0 on
 * error.  *shifted = 1 when the first pass needed the shift.  triangular = 0
 * allows the polar finish of the gauge move (R then a full matrix). */
int tn_cholqr_test(tn_ctx* ctx, int m, int n, const double* A, double tol, double* Q, double* R, int* shifted,
                   int triangular);

#ifdef __cplusplus
}
#endif
#endif
