/*
 * tn_oracle.h -- CPU oracle for the variational MPS ground-state sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (JanReimers/TensorNetworks, one-site DMRG) used as the checker for
 * the HIP path and as the timed CPU baseline.  Only tests/, the smoke() entry
 * and the cpu_baseline leg of bench.py may load it.  The product library
 * (tensornetworks_amd) never links or calls it.
 *
 * Arithmetic: real double.  The reference runs complex double (dcmplx) but
 * every operator on the path (Heisenberg/XXZ MPO, S+/S-/Sz) is real, so the
 * Hermitian eigenproblems are real-symmetric and the energies coincide.
 *
 * Layouts (all 0-based):
 *  reference site tensor   psi[m + d*(i1 + D1*i2)]          (Matrix6.H:110-111)
 *  reference environment   F[w + Dw*(i + D*j)], i=bra j=ket  (Vector3.H:99-102)
 *  operator-valued W       W[((w1*Dw2 + w2)*d + m)*d + n], m=bra n=ket
 *                                                           (OperatorValuedMatrix.H:11)
 *  internal (GEMM-ordered) layouts used by the matrix-free path and the GPU:
 *    site  A[a][m][b]      L-env  L[a][w][a']      R-env  R[w][b][b']
 *
 * Parity status of each function is stated next to it ("pinned by ...").
 */
#ifndef TN_ORACLE_H
#define TN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int orc_version(void);

/* ---- spin operators and MPO ------------------------------------------- */
/* SpinCalculator.C:20-70: index n <-> Sz = n - S; S+(m,n)!=0 iff m==n+1.
 * Pinned by the H12 string fixtures (ExactDiagTests.C:79-100) via tests. */
void orc_spin_ops(int d, double* Sp, double* Sm, double* Sz);

/* Hamiltonian_1D_NN_Heisenberg.C:35-79 (+ Dz*Sz^2 added to the on-site slot,
 * an extension used by the S=1 workload).  form: 0=RegularUpper, 1=RegularLower.
 * Wbulk is 5x5 blocks of dxd.  Pinned by closed forms E0(L=2,S)=-S(S+1)
 * and ED energies (ExactDiagTests.C:172-301). */
void orc_xxz_bulk_W(int d, double Jxy, double Jz, double hz, double Dz, int form,
                    double* Wbulk);
/* SiteOperatorImp.C:109-134 Init_lr: edge rows/cols.  pos: 0=left edge,
 * 1=bulk, 2=right edge, 3=single site.  Returns Dw1,Dw2 and writes W. */
void orc_site_W(int d, int Dw, const double* Wbulk, int form, int pos,
                double* W, int* Dw1, int* Dw2);
/* Product MPO for <H^2>: W2[(w1,v1),(w2,v2)](m,n) = sum_k W(w1,w2)(m,k) W(v1,v2)(k,n)
 * (OperatorValuedMatrix.H:410 TensorProduct, uncompressed). */
void orc_mpo_product(int d, int Dw1, int Dw2, const double* Wa, int Ew1, int Ew2,
                     const double* Wb, double* W2);

/* ---- contractions in the reference layout ------------------------------ */
/* MPSSite_Contractions.C:155-193 GetHeff: dense N x N, N=d*D1*D2, row-major
 * H[row*N+col], row=(m,i1,i2) bra, col=(n,j1,j2) ket.  Pinned by Hermiticity
 * and by the whole-sweep golden energies. */
void orc_heff_dense(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                    const double* L, const double* R, double* H);
/* Matrix-free y = Heff x in the reference layout (same sum as GetHeff). */
void orc_heff_apply_ref(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                        const double* L, const double* R, const double* x, double* y);
/* MPSSite_Contractions.C:292-348 IterateLeft_F and :352-400 IterateRightF.
 * Pinned by Neel closed form, env closure L(L+1)=R(0)=<H> (MPOTests.C:115-275). */
void orc_env_left_ref(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                      const double* Fin, const double* M, double* Fout);
void orc_env_right_ref(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                       const double* Fin, const double* M, double* Fout);

/* ---- layout conversions (reference <-> internal) ----------------------- */
void orc_psi_ref_to_int(int d, int D1, int D2, const double* ref, double* in);
void orc_psi_int_to_ref(int d, int D1, int D2, const double* in, double* ref);
void orc_lenv_ref_to_int(int Dw, int D, const double* ref, double* in);
void orc_lenv_int_to_ref(int Dw, int D, const double* in, double* ref);
void orc_renv_ref_to_int(int Dw, int D, const double* ref, double* in);
void orc_renv_int_to_ref(int Dw, int D, const double* in, double* ref);

/* ---- GEMM-ordered matrix-free path (internal layout, OpenMP) ----------- */
/* Same sums as the reference functions above, in optimal contraction order;
 * the timed CPU baseline.  Pinned against the *_ref functions in tests. */
void orc_gemm(int M, int N, int K, const double* A, long sai, long sak,
              const double* B, long sbk, long sbj, double beta, double* C, long ldc);
void orc_heff_apply_int(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                        const double* L, const double* R, const double* x, double* y,
                        double* work);
long orc_heff_work_size(int d, int D1, int D2, int Dw1, int Dw2);
void orc_env_left_int(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                      const double* Lin, const double* A, double* Lout, double* work);
void orc_env_right_int(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                       const double* Rin, const double* A, double* Rout, double* work);

/* ---- dense linear algebra ---------------------------------------------- */
/* Cyclic Jacobi, all eigenpairs ascending; Z column-major (column j = vector j).
 * Stands in for zheevx (LapackEigenSolver.C:62-109).  Pinned by the EigenTester
 * invariants (LinearAlgebraTests.H:206-218) and numpy.linalg.eigh in tests. */
int orc_syev(int n, const double* A, double* w, double* Z);
/* One-sided Jacobi thin SVD of column-major A (m x n): U (m x k col-major),
 * s (k, descending), Vt (k x n col-major), k=min(m,n), U/Vt orthonormal even
 * for zero singular values.  Stands in for zgesvd (LapackSVDSolver.C:42-90).
 * Pinned by reconstruction/orthonormality and numpy.linalg.svd in tests. */
int orc_svd(int m, int n, const double* A, double* U, double* s, double* Vt);
/* SVCompressorImp.C:19-73: returns kept D; *discarded = sum of dropped s^2;
 * *rescale = squared-norm ratio applied to kept s.  Pinned by closed-form cases. */
int orc_svcompress(int N, const double* s, int Dmax, double eps, double* discarded,
                   double* rescale);

/* ---- Lanczos (lowest eigenpair), full re-orthogonalisation -------------- */
typedef void (*orc_matvec_fn)(void* ctx, const double* x, double* y);
/* Returns number of matvecs; x0 in/out (start vector / Ritz vector). */
int orc_lanczos(long N, orc_matvec_fn mv, void* ctx, double* x0, double tol, int maxit,
                double* e0, double* e1);

/* ---- exact diagonalisation (Lanczos on the d^L space) ------------------ */
/* H = sum_bonds [Jxy/2(S+S- + S-S+) + Jz SzSz] + sum_sites (hz Sz + Dz Sz^2),
 * the MPO convention (hz once per site).  FullStateImp.C:91-133 pattern.
 * Pinned by ED energies in ExactDiagTests.C:186-301 and test_Unittest1.py:64-66. */
double orc_ed_ground(int L, int d, double Jxy, double Jz, double hz, double Dz,
                     double tol, int maxit);

/* Ground energy of any MPO (given like the chain's: dw1, dw2, W concatenated)
 * by Lanczos on the d^L space, the MPO applied site by site.  Pinned against
 * orc_ed_ground on the XXZ chain, numpy eigvalsh of the contracted MPO, the
 * free-fermion closed form of the transverse-field Ising chain and brute force
 * over classical configurations (tests/test_oracle.py, tests/test_mpo.py). */
double orc_ed_ground_mpo(int L, int d, const int* dw1, const int* dw2, const double* wcat, double tol,
                         int maxit);

/* ---- the sweep (MPSImp_Variational.C:16-173) ---------------------------- */
typedef struct orc_mps orc_mps;
/* MPSImp.C:69-90 shapes; init 0=Product 1=Random(seeded U(-1,1)) 2=Neel. */
orc_mps* orc_mps_create(int L, int d, double Jxy, double Jz, double hz, double Dz,
                        int form, int D, int init, uint64_t seed);
void orc_mps_destroy(orc_mps* s);
/* mode: 0 = reference-faithful (dense Heff + full Jacobi eigensolve),
 *       1 = matrix-free (GEMM-ordered Heff + Lanczos to tol eps_eig). */
void orc_mps_normalize(orc_mps* s, int dir); /* dir 0=DLeft 1=DRight */
void orc_mps_load_caches(orc_mps* s);
/* One schedule line; returns <H^2>-<H>^2, *nsweeps = sweeps of the last D. */
double orc_mps_find_gs_line(orc_mps* s, int maxiter, int Dmax, int DeltaD,
                            double epsE, double epsL, double epsEig, int mode,
                            int* nsweeps);
void orc_mps_sweep(orc_mps* s, int dir, double epsEig, int mode);
double orc_mps_expectation(orc_mps* s, int which); /* 0=H, 1=H^2, 2=identity */
int orc_mps_get_dims(const orc_mps* s, int* D1, int* D2);
int orc_mps_get_site(const orc_mps* s, int i, double* M); /* reference layout */
int orc_mps_set_site(orc_mps* s, int i, int D1, int D2, const double* M);
double orc_mps_max_delta_e(const orc_mps* s);
double orc_mps_max_delta_l(const orc_mps* s);
int orc_mps_total_matvecs(const orc_mps* s);
void orc_mps_norm_status(const orc_mps* s, char* out); /* L chars + NUL */
int orc_mps_get_bond_sv(const orc_mps* s, int b, double* sv);
/* two-site sweep (extension; no reference counterpart): one pass in dir with
 * SVCompressor rule (Dmax, epsSV); returns max truncation error. */
double orc_mps_sweep2(orc_mps* s, int dir, int Dmax, double epsSV, double epsEig);
double orc_mps_last_energy(const orc_mps* s);

/* ---- learned Davidson preconditioner (no reference counterpart) -------
 * The network maps per-element features of a local eigenproblem to a factor
 * g on the Davidson diagonal correction: t_i = g_i * s0_i * r_i with
 * s0_i = -1/den_i, den_i = diag_i - theta (|den_i| < delta -> delta).
 * Features (raw): x = [asinh((diag_i - theta)/delta), (theta - mean(diag))/sd(diag),
 * log(lam_left[a] + 1e-10), log(lam_right[b] + 1e-10)] (mean and sd over the n
 * given elements), standardised by
 * norm = [mu0..mu3, inv_sd0..inv_sd3].  Network 4 -> 64 -> 64 -> 1, tanh,
 * params (ORC_MLP_NP) = W1[64][4], b1[64], W2[64][64], b2[64], w3[64], b3.
 * Loss over n samples: (1/n) sum (g_i a_i - t_i)^2 (a_i = s0_i r_i, t_i the
 * exact correction).  Pinned by central finite differences of the loss
 * (tests/test_oracle.py) and used as the checker of the GPU kernels. */
#define ORC_MLP_H 64
#define ORC_MLP_NP (ORC_MLP_H * 4 + ORC_MLP_H + ORC_MLP_H * ORC_MLP_H + ORC_MLP_H + ORC_MLP_H + 1)
void orc_mlp_features(long n, const double* diag, double theta, const double* lam_l, const double* lam_r,
                      double delta, double* X);
void orc_mlp_forward(const double* params, const double* norm, long n, const double* X, double* g);
double orc_mlp_loss_grad(const double* params, const double* norm, long n, const double* X, const double* a,
                         const double* t, double* grad);

/* ---- seeded generator shared with the GPU host code ------------------- */
uint64_t orc_rng_next(uint64_t* state);
double orc_rng_uniform(uint64_t* state); /* U(-1,1) */

#ifdef __cplusplus
}
#endif
#endif
