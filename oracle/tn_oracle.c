/*
 * tn_oracle.c -- CPU oracle (test infrastructure, see tn_oracle.h).
 *
 * A real-double C restatement of JanReimers/TensorNetworks' one-site
 * variational sweep.  Each function cites the reference file:line it follows.
 * Nothing here is shipped with or called by the HIP product path.
 */
#include "tn_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_version(void) { return 1; }

static void* xcalloc(size_t n, size_t sz)
{
    void* p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "oracle: out of memory (%zu x %zu)\n", n, sz); abort(); }
    return p;
}
static double* dalloc(size_t n) { return (double*)xcalloc(n, sizeof(double)); }
static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* seeded generator: splitmix64, U(-1,1) from the top 53 bits                */
uint64_t orc_rng_next(uint64_t* s)
{
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
double orc_rng_uniform(uint64_t* s)
{
    return (double)(orc_rng_next(s) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

/* ------------------------------------------------------------------------ */
/* SpinCalculator.C:20-70                                                    */
void orc_spin_ops(int d, double* Sp, double* Sm, double* Sz)
{
    double S = 0.5 * (d - 1);
    for (int m = 0; m < d; m++)
        for (int n = 0; n < d; n++) {
            double sm = m - S, sn = n - S;
            Sp[m * d + n] = (m == n + 1) ? sqrt(S * (S + 1.0) - sm * sn) : 0.0;
            Sm[m * d + n] = (m + 1 == n) ? sqrt(S * (S + 1.0) - sm * sn) : 0.0;
            Sz[m * d + n] = (m == n) ? sm : 0.0;
        }
}

/* Hamiltonian_1D_NN_Heisenberg.C:35-79 (Dz*Sz^2 joins hz*Sz in the on-site slot) */
void orc_xxz_bulk_W(int d, double Jxy, double Jz, double hz, double Dz, int form,
                    double* W)
{
    const int Dw = 5, dd = d * d;
    double *Sp = dalloc(dd), *Sm = dalloc(dd), *Sz = dalloc(dd), *I = dalloc(dd),
           *On = dalloc(dd);
    orc_spin_ops(d, Sp, Sm, Sz);
    for (int m = 0; m < d; m++) {
        I[m * d + m] = 1.0;
        On[m * d + m] = hz * Sz[m * d + m] + Dz * Sz[m * d + m] * Sz[m * d + m];
    }
    memset(W, 0, sizeof(double) * Dw * Dw * dd);
#define WBLK(a, b) (W + ((a) * Dw + (b)) * dd)
    for (int k = 0; k < dd; k++) {
        if (form == 0) { /* RegularUpper */
            WBLK(0, 0)[k] = I[k];
            WBLK(0, 1)[k] = Sp[k];
            WBLK(0, 2)[k] = Sm[k];
            WBLK(0, 3)[k] = Sz[k];
            WBLK(0, 4)[k] = On[k];
            WBLK(1, 4)[k] = 0.5 * Jxy * Sm[k];
            WBLK(2, 4)[k] = 0.5 * Jxy * Sp[k];
            WBLK(3, 4)[k] = Jz * Sz[k];
            WBLK(4, 4)[k] = I[k];
        } else { /* RegularLower */
            WBLK(0, 0)[k] = I[k];
            WBLK(1, 0)[k] = Sp[k];
            WBLK(2, 0)[k] = Sm[k];
            WBLK(3, 0)[k] = Sz[k];
            WBLK(4, 0)[k] = On[k];
            WBLK(4, 1)[k] = 0.5 * Jxy * Sm[k];
            WBLK(4, 2)[k] = 0.5 * Jxy * Sp[k];
            WBLK(4, 3)[k] = Jz * Sz[k];
            WBLK(4, 4)[k] = I[k];
        }
    }
#undef WBLK
    free(Sp); free(Sm); free(Sz); free(I); free(On);
}

/* SiteOperatorImp.C:33-51 (lindex/rindex per form) and :109-134 Init_lr */
void orc_site_W(int d, int Dw, const double* Wb, int form, int pos, double* W,
                int* Dw1, int* Dw2)
{
    const int dd = d * d;
    int lindex = (form == 0) ? 0 : Dw - 1;
    int rindex = (form == 0) ? Dw - 1 : 0;
    int r0 = 0, r1 = Dw, c0 = 0, c1 = Dw;
    if (pos == 0 || pos == 3) { r0 = lindex; r1 = lindex + 1; }
    if (pos == 2 || pos == 3) { c0 = rindex; c1 = rindex + 1; }
    *Dw1 = r1 - r0;
    *Dw2 = c1 - c0;
    for (int a = r0; a < r1; a++)
        for (int b = c0; b < c1; b++)
            memcpy(W + ((a - r0) * (*Dw2) + (b - c0)) * dd, Wb + (a * Dw + b) * dd,
                   sizeof(double) * dd);
}

/* OperatorValuedMatrix.H:410 TensorProduct of two W's (uncompressed H^2 MPO) */
void orc_mpo_product(int d, int Dw1, int Dw2, const double* Wa, int Ew1, int Ew2,
                     const double* Wb, double* W2)
{
    const int dd = d * d, F2 = Dw2 * Ew2;
    for (int w1 = 0; w1 < Dw1; w1++)
        for (int v1 = 0; v1 < Ew1; v1++)
            for (int w2 = 0; w2 < Dw2; w2++)
                for (int v2 = 0; v2 < Ew2; v2++) {
                    const double* A = Wa + (w1 * Dw2 + w2) * dd;
                    const double* B = Wb + (v1 * Ew2 + v2) * dd;
                    double* C = W2 + ((w1 * Ew1 + v1) * F2 + (w2 * Ew2 + v2)) * dd;
                    for (int m = 0; m < d; m++)
                        for (int n = 0; n < d; n++) {
                            double acc = 0;
                            for (int k = 0; k < d; k++) acc += A[m * d + k] * B[k * d + n];
                            C[m * d + n] = acc;
                        }
                }
}

/* ------------------------------------------------------------------------ */
/* reference-layout contractions                                             */
#define PSI(m, i1, i2) ((m) + d * ((i1) + D1 * (i2)))
#define ENV(w, i, j, Dw, D) ((w) + (Dw) * ((i) + (D) * (j)))
#define WEL(w1, w2, m, n) W[((((w1) * Dw2 + (w2)) * d + (m)) * d + (n))]

/* MPSSite_Contractions.C:155-193 */
void orc_heff_dense(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                    const double* L, const double* R, double* H)
{
    const long N = (long)d * D1 * D2;
    double* WR = dalloc(Dw1);
    for (int m = 0; m < d; m++)
        for (int n = 0; n < d; n++)
            for (int i2 = 0; i2 < D2; i2++)
                for (int j2 = 0; j2 < D2; j2++) {
                    for (int w1 = 0; w1 < Dw1; w1++) {
                        double acc = 0;
                        for (int w2 = 0; w2 < Dw2; w2++)
                            if (WEL(w1, w2, m, n) != 0.0)
                                acc += WEL(w1, w2, m, n) * R[ENV(w2, i2, j2, Dw2, D2)];
                        WR[w1] = acc;
                    }
                    for (int i1 = 0; i1 < D1; i1++)
                        for (int j1 = 0; j1 < D1; j1++) {
                            double lwr = 0;
                            for (int w1 = 0; w1 < Dw1; w1++)
                                lwr += L[ENV(w1, i1, j1, Dw1, D1)] * WR[w1];
                            H[PSI(m, i1, i2) * N + PSI(n, j1, j2)] = lwr;
                        }
                }
    free(WR);
}

void orc_heff_apply_ref(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                        const double* L, const double* R, const double* x, double* y)
{
    /* T(w1,m,j1,i2) = sum_{w2,n,j2} W(w1,w2)(m,n) R(w2,i2,j2) x(n,j1,j2) */
    double* T = dalloc((size_t)Dw1 * d * D1 * D2);
#define TIX(w1, m, j1, i2) ((((size_t)(w1) * d + (m)) * D1 + (j1)) * D2 + (i2))
    for (int w1 = 0; w1 < Dw1; w1++)
        for (int m = 0; m < d; m++)
            for (int j1 = 0; j1 < D1; j1++)
                for (int i2 = 0; i2 < D2; i2++) {
                    double acc = 0;
                    for (int w2 = 0; w2 < Dw2; w2++)
                        for (int n = 0; n < d; n++) {
                            double wv = WEL(w1, w2, m, n);
                            if (wv == 0.0) continue;
                            double s = 0;
                            for (int j2 = 0; j2 < D2; j2++)
                                s += R[ENV(w2, i2, j2, Dw2, D2)] * x[PSI(n, j1, j2)];
                            acc += wv * s;
                        }
                    T[TIX(w1, m, j1, i2)] = acc;
                }
    for (int m = 0; m < d; m++)
        for (int i1 = 0; i1 < D1; i1++)
            for (int i2 = 0; i2 < D2; i2++) {
                double acc = 0;
                for (int w1 = 0; w1 < Dw1; w1++)
                    for (int j1 = 0; j1 < D1; j1++)
                        acc += L[ENV(w1, i1, j1, Dw1, D1)] * T[TIX(w1, m, j1, i2)];
                y[PSI(m, i1, i2)] = acc;
            }
#undef TIX
    free(T);
}

/* MPSSite_Contractions.C:292-348  IterateLeft_F:
 * F(w2,i2,j2) = sum_{m,i1} M[m](i1,i2) sum_n sum_w1 W(w1,w2)(m,n) sum_j1 Fin(w1,i1,j1) M[n](j1,j2) */
void orc_env_left_ref(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                      const double* Fin, const double* M, double* Fout)
{
    /* FA(w1,n,i1,j2) = sum_j1 Fin(w1,i1,j1) M[n](j1,j2)   (ContractFA) */
    double* FA = dalloc((size_t)Dw1 * d * D1 * D2);
#define FAX(w1, n, i1, j2) ((((size_t)(w1) * d + (n)) * D1 + (i1)) * D2 + (j2))
    for (int w1 = 0; w1 < Dw1; w1++)
        for (int n = 0; n < d; n++)
            for (int i1 = 0; i1 < D1; i1++)
                for (int j2 = 0; j2 < D2; j2++) {
                    double acc = 0;
                    for (int j1 = 0; j1 < D1; j1++)
                        acc += Fin[ENV(w1, i1, j1, Dw1, D1)] * M[PSI(n, j1, j2)];
                    FA[FAX(w1, n, i1, j2)] = acc;
                }
    for (int w2 = 0; w2 < Dw2; w2++)
        for (int i2 = 0; i2 < D2; i2++)
            for (int j2 = 0; j2 < D2; j2++) {
                double awfa = 0;
                for (int m = 0; m < d; m++)
                    for (int i1 = 0; i1 < D1; i1++) {
                        double wfa = 0; /* ContractWFA */
                        for (int n = 0; n < d; n++)
                            for (int w1 = 0; w1 < Dw1; w1++) {
                                double wv = WEL(w1, w2, m, n);
                                if (wv != 0.0) wfa += wv * FA[FAX(w1, n, i1, j2)];
                            }
                        awfa += M[PSI(m, i1, i2)] * wfa;
                    }
                Fout[ENV(w2, i2, j2, Dw2, D2)] = awfa;
            }
#undef FAX
    free(FA);
}

/* MPSSite_Contractions.C:352-400  IterateRightF:
 * F(w1,i1,j1) = sum_{m,i2} M[m](i1,i2) sum_n sum_w2 W(w1,w2)(m,n) sum_j2 Fin(w2,i2,j2) M[n](j1,j2) */
void orc_env_right_ref(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                       const double* Fin, const double* M, double* Fout)
{
    double* FB = dalloc((size_t)Dw2 * d * D2 * D1);
#define FBX(w2, n, i2, j1) ((((size_t)(w2) * d + (n)) * D2 + (i2)) * D1 + (j1))
    for (int w2 = 0; w2 < Dw2; w2++)
        for (int n = 0; n < d; n++)
            for (int i2 = 0; i2 < D2; i2++)
                for (int j1 = 0; j1 < D1; j1++) {
                    double acc = 0;
                    for (int j2 = 0; j2 < D2; j2++)
                        acc += Fin[ENV(w2, i2, j2, Dw2, D2)] * M[PSI(n, j1, j2)];
                    FB[FBX(w2, n, i2, j1)] = acc;
                }
    for (int w1 = 0; w1 < Dw1; w1++)
        for (int i1 = 0; i1 < D1; i1++)
            for (int j1 = 0; j1 < D1; j1++) {
                double bwfb = 0;
                for (int m = 0; m < d; m++)
                    for (int i2 = 0; i2 < D2; i2++) {
                        double wfb = 0;
                        for (int n = 0; n < d; n++)
                            for (int w2 = 0; w2 < Dw2; w2++) {
                                double wv = WEL(w1, w2, m, n);
                                if (wv != 0.0) wfb += wv * FB[FBX(w2, n, i2, j1)];
                            }
                        bwfb += M[PSI(m, i1, i2)] * wfb;
                    }
                Fout[ENV(w1, i1, j1, Dw1, D1)] = bwfb;
            }
#undef FBX
    free(FB);
}

/* ------------------------------------------------------------------------ */
/* layout conversions                                                        */
void orc_psi_ref_to_int(int d, int D1, int D2, const double* ref, double* in)
{
    for (int i1 = 0; i1 < D1; i1++)
        for (int m = 0; m < d; m++)
            for (int i2 = 0; i2 < D2; i2++)
                in[((size_t)i1 * d + m) * D2 + i2] = ref[PSI(m, i1, i2)];
}
void orc_psi_int_to_ref(int d, int D1, int D2, const double* in, double* ref)
{
    for (int i1 = 0; i1 < D1; i1++)
        for (int m = 0; m < d; m++)
            for (int i2 = 0; i2 < D2; i2++)
                ref[PSI(m, i1, i2)] = in[((size_t)i1 * d + m) * D2 + i2];
}
void orc_lenv_ref_to_int(int Dw, int D, const double* ref, double* in)
{
    for (int i = 0; i < D; i++)
        for (int w = 0; w < Dw; w++)
            for (int j = 0; j < D; j++)
                in[((size_t)i * Dw + w) * D + j] = ref[ENV(w, i, j, Dw, D)];
}
void orc_lenv_int_to_ref(int Dw, int D, const double* in, double* ref)
{
    for (int i = 0; i < D; i++)
        for (int w = 0; w < Dw; w++)
            for (int j = 0; j < D; j++)
                ref[ENV(w, i, j, Dw, D)] = in[((size_t)i * Dw + w) * D + j];
}
void orc_renv_ref_to_int(int Dw, int D, const double* ref, double* in)
{
    for (int w = 0; w < Dw; w++)
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++)
                in[((size_t)w * D + i) * D + j] = ref[ENV(w, i, j, Dw, D)];
}
void orc_renv_int_to_ref(int Dw, int D, const double* in, double* ref)
{
    for (int w = 0; w < Dw; w++)
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++)
                ref[ENV(w, i, j, Dw, D)] = in[((size_t)w * D + i) * D + j];
}

/* ------------------------------------------------------------------------ */
/* blocked GEMM (packed panels, 4x8 register tile, OpenMP over C tiles)      */
enum { GMC = 64, GNC = 256, GKC = 256 };

void orc_gemm(int M, int N, int K, const double* A, long sai, long sak,
              const double* B, long sbk, long sbj, double beta, double* C, long ldc)
{
    if (M <= 0 || N <= 0) return;
    if (beta != 1.0) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < M; i++)
            for (int j = 0; j < N; j++)
                C[(long)i * ldc + j] = (beta == 0.0) ? 0.0 : beta * C[(long)i * ldc + j];
    }
    if (K <= 0) return;
    const int nbi = (M + GMC - 1) / GMC, nbj = (N + GNC - 1) / GNC;
#pragma omp parallel
    {
        double* Ap = (double*)aligned_alloc(64, sizeof(double) * GMC * GKC);
        double* Bp = (double*)aligned_alloc(64, sizeof(double) * GKC * GNC);
#pragma omp for schedule(dynamic) collapse(2)
        for (int bi = 0; bi < nbi; bi++)
            for (int bj = 0; bj < nbj; bj++) {
                const int i0 = bi * GMC, j0 = bj * GNC;
                const int mc = imin(GMC, M - i0), nc = imin(GNC, N - j0);
                const int mcp = (mc + 3) & ~3, ncp = (nc + 7) & ~7;
                for (int k0 = 0; k0 < K; k0 += GKC) {
                    const int kc = imin(GKC, K - k0);
                    for (int i = 0; i < mcp; i++)
                        for (int k = 0; k < kc; k++)
                            Ap[i * GKC + k] = (i < mc) ? A[(long)(i0 + i) * sai + (long)(k0 + k) * sak] : 0.0;
                    for (int k = 0; k < kc; k++)
                        for (int j = 0; j < ncp; j++)
                            Bp[k * GNC + j] = (j < nc) ? B[(long)(k0 + k) * sbk + (long)(j0 + j) * sbj] : 0.0;
                    for (int ii = 0; ii < mcp; ii += 4)
                        for (int jj = 0; jj < ncp; jj += 8) {
                            double acc[4][8];
                            memset(acc, 0, sizeof(acc));
                            for (int k = 0; k < kc; k++) {
                                const double* b = Bp + k * GNC + jj;
                                for (int r = 0; r < 4; r++) {
                                    const double a = Ap[(ii + r) * GKC + k];
                                    for (int c = 0; c < 8; c++) acc[r][c] += a * b[c];
                                }
                            }
                            for (int r = 0; r < 4; r++) {
                                if (ii + r >= mc) break;
                                double* cr = C + (long)(i0 + ii + r) * ldc + j0 + jj;
                                const int cn = imin(8, nc - jj);
                                for (int c = 0; c < cn; c++) cr[c] += acc[r][c];
                            }
                        }
                }
            }
        free(Ap);
        free(Bp);
    }
}

/* ------------------------------------------------------------------------ */
/* internal-layout matrix-free path (site A[a][m][b], L[a][w][a'], R[w][b][b']) */
long orc_heff_work_size(int d, int D1, int D2, int Dw1, int Dw2)
{
    long t1 = (long)D1 * d * Dw2 * D2, t2 = (long)Dw1 * D1 * d * D2;
    long e1 = (long)D1 * Dw1 * d * D2, e2 = (long)D2 * Dw2 * d * D1; /* env temps */
    long a = t1 > e1 ? t1 : e1, b = t2 > e2 ? t2 : e2;
    a = a > e2 ? a : e2;
    b = b > e1 ? b : e1;
    return a + b;
}

/* y[a][m][b] = sum L[a][w][a'] W[w][w'][m][n] R[w'][b][b'] x[a'][n][b'] */
void orc_heff_apply_int(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                        const double* L, const double* R, const double* x, double* y,
                        double* work)
{
    double* T1 = work;                                 /* [a'][n][w'][b] */
    double* T2 = work + (long)D1 * d * Dw2 * D2;       /* [w][a'][m][b]  */
    /* S1: T1[(a',n)][(w',b)] = sum_b' x[(a',n)][b'] R[(w',b)][b'] */
    orc_gemm(D1 * d, Dw2 * D2, D2, x, D2, 1, R, 1, D2, 0.0, T1, (long)Dw2 * D2);
    /* S2: T2[w][a'][m][b] = sum_{w',n} W[w][w'][m][n] T1[a'][n][w'][b] */
#pragma omp parallel for collapse(2) schedule(static)
    for (int w = 0; w < Dw1; w++)
        for (int ap = 0; ap < D1; ap++)
            for (int m = 0; m < d; m++) {
                double* out = T2 + (((long)w * D1 + ap) * d + m) * D2;
                for (int b = 0; b < D2; b++) out[b] = 0.0;
                for (int wp = 0; wp < Dw2; wp++)
                    for (int n = 0; n < d; n++) {
                        double wv = W[(((long)w * Dw2 + wp) * d + m) * d + n];
                        if (wv == 0.0) continue;
                        const double* in = T1 + (((long)ap * d + n) * Dw2 + wp) * D2;
                        for (int b = 0; b < D2; b++) out[b] += wv * in[b];
                    }
            }
    /* S3: y[a][(m,b)] = sum_{(w,a')} L[a][(w,a')] T2[(w,a')][(m,b)] */
    orc_gemm(D1, d * D2, Dw1 * D1, L, (long)Dw1 * D1, 1, T2, (long)d * D2, 1, 0.0, y,
             (long)d * D2);
}

/* L'[b][w'][b'] = sum A[a][m][b] W[w][w'][m][n] L[a][w][a'] A[a'][n][b'] */
void orc_env_left_int(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                      const double* Lin, const double* A, double* Lout, double* work)
{
    double* X = work;                                  /* [a][w][n][b'] */
    double* Y = work + (long)D1 * Dw1 * d * D2;        /* [a][m][w'][b'] */
    orc_gemm(D1 * Dw1, d * D2, D1, Lin, D1, 1, A, (long)d * D2, 1, 0.0, X, (long)d * D2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int a = 0; a < D1; a++)
        for (int m = 0; m < d; m++)
            for (int wp = 0; wp < Dw2; wp++) {
                double* out = Y + (((long)a * d + m) * Dw2 + wp) * D2;
                for (int b = 0; b < D2; b++) out[b] = 0.0;
                for (int w = 0; w < Dw1; w++)
                    for (int n = 0; n < d; n++) {
                        double wv = W[(((long)w * Dw2 + wp) * d + m) * d + n];
                        if (wv == 0.0) continue;
                        const double* in = X + (((long)a * Dw1 + w) * d + n) * D2;
                        for (int b = 0; b < D2; b++) out[b] += wv * in[b];
                    }
            }
    /* L'[b][(w',b')] = sum_{(a,m)} A[(a,m)][b] Y[(a,m)][(w',b')] */
    orc_gemm(D2, Dw2 * D2, D1 * d, A, 1, D2, Y, (long)Dw2 * D2, 1, 0.0, Lout, (long)Dw2 * D2);
}

/* R'[w][a][a'] = sum A[a][m][b] W[w][w'][m][n] R[w'][b][b'] A[a'][n][b'] */
void orc_env_right_int(int d, int D1, int D2, int Dw1, int Dw2, const double* W,
                       const double* Rin, const double* A, double* Rout, double* work)
{
    double* X = work;                                  /* [a'][n][w'][b] */
    double* Y = work + (long)D1 * d * Dw2 * D2;        /* [a'][w][m][b]  */
    orc_gemm(D1 * d, Dw2 * D2, D2, A, D2, 1, Rin, 1, D2, 0.0, X, (long)Dw2 * D2);
#pragma omp parallel for collapse(2) schedule(static)
    for (int ap = 0; ap < D1; ap++)
        for (int w = 0; w < Dw1; w++)
            for (int m = 0; m < d; m++) {
                double* out = Y + (((long)ap * Dw1 + w) * d + m) * D2;
                for (int b = 0; b < D2; b++) out[b] = 0.0;
                for (int wp = 0; wp < Dw2; wp++)
                    for (int n = 0; n < d; n++) {
                        double wv = W[(((long)w * Dw2 + wp) * d + m) * d + n];
                        if (wv == 0.0) continue;
                        const double* in = X + (((long)ap * d + n) * Dw2 + wp) * D2;
                        for (int b = 0; b < D2; b++) out[b] += wv * in[b];
                    }
            }
    /* R'[w][a][a'] = sum_{(m,b)} A[a][(m,b)] Y[a'][w][(m,b)] */
    for (int w = 0; w < Dw1; w++)
        orc_gemm(D1, D1, d * D2, A, (long)d * D2, 1, Y + (long)w * d * D2, 1,
                 (long)Dw1 * d * D2, 0.0, Rout + (long)w * D1 * D1, D1);
}

/* ------------------------------------------------------------------------ */
/* cyclic Jacobi eigensolver (stands in for zheevx, LapackEigenSolver.C:62-109) */
int orc_syev(int n, const double* A0, double* w, double* Z)
{
    double* A = dalloc((size_t)n * n);
    memcpy(A, A0, sizeof(double) * n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Z[(long)j * n + i] = (i == j);
    int sweep;
    for (sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double v = A[(long)i * n + j] * A[(long)i * n + j];
                tot += v;
                if (i != j) off += v;
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[(long)p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double app = A[(long)p * n + p], aqq = A[(long)q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) { /* columns p,q */
                    double akp = A[(long)k * n + p], akq = A[(long)k * n + q];
                    A[(long)k * n + p] = c * akp - s * akq;
                    A[(long)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) { /* rows p,q */
                    double apk = A[(long)p * n + k], aqk = A[(long)q * n + k];
                    A[(long)p * n + k] = c * apk - s * aqk;
                    A[(long)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double zkp = Z[(long)p * n + k], zkq = Z[(long)q * n + k];
                    Z[(long)p * n + k] = c * zkp - s * zkq;
                    Z[(long)q * n + k] = s * zkp + c * zkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = A[(long)i * n + i];
    /* sort ascending (selection sort; n is small) */
    double* tmp = dalloc(n);
    for (int i = 0; i < n - 1; i++) {
        int k = i;
        for (int j = i + 1; j < n; j++)
            if (w[j] < w[k]) k = j;
        if (k != i) {
            double t = w[i]; w[i] = w[k]; w[k] = t;
            memcpy(tmp, Z + (long)i * n, sizeof(double) * n);
            memcpy(Z + (long)i * n, Z + (long)k * n, sizeof(double) * n);
            memcpy(Z + (long)k * n, tmp, sizeof(double) * n);
        }
    }
    free(tmp);
    free(A);
    return sweep;
}

/* ------------------------------------------------------------------------ */
/* one-sided Jacobi SVD (stands in for zgesvd, LapackSVDSolver.C:42-90)       */
static double dot(long n, const double* a, const double* b)
{
    double s = 0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* orthonormalise column j of Q (m rows, column-major) against columns
 * listed in 'keep'; fills with a unit vector if needed. */
static void complete_column(int m, int ncols, double* Q, const int* good, int j)
{
    double* q = Q + (long)j * m;
    for (int e = 0; e < m + 1; e++) {
        for (int i = 0; i < m; i++) q[i] = (i == (e + j) % m) ? 1.0 : 0.0;
        for (int pass = 0; pass < 2; pass++)
            for (int k = 0; k < ncols; k++) {
                if (k == j || !good[k]) continue;
                double* o = Q + (long)k * m;
                double c = dot(m, o, q);
                for (int i = 0; i < m; i++) q[i] -= c * o[i];
            }
        double nr = sqrt(dot(m, q, q));
        if (nr > 0.5) {
            for (int i = 0; i < m; i++) q[i] /= nr;
            return;
        }
    }
}

static void svd_tall(int m, int n, const double* A, double* U, double* s, double* V)
{
    /* m >= n.  U (m x n) <- A, V (n x n) <- I, rotate columns until orthogonal. */
    memcpy(U, A, sizeof(double) * m * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[(long)j * n + i] = (i == j);
    for (int sweep = 0; sweep < 80; sweep++) {
        int rotated = 0;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double* up = U + (long)p * m;
                double* uq = U + (long)q * m;
                double alpha = dot(m, up, up), beta = dot(m, uq, uq), gamma = dot(m, up, uq);
                if (fabs(gamma) <= 1e-15 * sqrt(alpha * beta) || gamma == 0.0) continue;
                rotated = 1;
                double zeta = (beta - alpha) / (2.0 * gamma);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                for (int i = 0; i < m; i++) {
                    double a = up[i], b = uq[i];
                    up[i] = c * a - sn * b;
                    uq[i] = sn * a + c * b;
                }
                double* vp = V + (long)p * n;
                double* vq = V + (long)q * n;
                for (int i = 0; i < n; i++) {
                    double a = vp[i], b = vq[i];
                    vp[i] = c * a - sn * b;
                    vq[i] = sn * a + c * b;
                }
            }
        if (!rotated) break;
    }
    for (int j = 0; j < n; j++) s[j] = sqrt(dot(m, U + (long)j * m, U + (long)j * m));
    /* sort descending */
    double* tmp = dalloc(m > n ? m : n);
    for (int i = 0; i < n - 1; i++) {
        int k = i;
        for (int j = i + 1; j < n; j++)
            if (s[j] > s[k]) k = j;
        if (k != i) {
            double t = s[i]; s[i] = s[k]; s[k] = t;
            memcpy(tmp, U + (long)i * m, sizeof(double) * m);
            memcpy(U + (long)i * m, U + (long)k * m, sizeof(double) * m);
            memcpy(U + (long)k * m, tmp, sizeof(double) * m);
            memcpy(tmp, V + (long)i * n, sizeof(double) * n);
            memcpy(V + (long)i * n, V + (long)k * n, sizeof(double) * n);
            memcpy(V + (long)k * n, tmp, sizeof(double) * n);
        }
    }
    free(tmp);
    int* good = (int*)xcalloc(n, sizeof(int));
    double smax = n > 0 ? s[0] : 0.0;
    for (int j = 0; j < n; j++) {
        if (s[j] > 1e-300 && s[j] > smax * 1e-15) {
            for (int i = 0; i < m; i++) U[(long)j * m + i] /= s[j];
            good[j] = 1;
        }
    }
    for (int j = 0; j < n; j++)
        if (!good[j]) {
            complete_column(m, n, U, good, j);
            good[j] = 1;
        }
    free(good);
}

int orc_svd(int m, int n, const double* A, double* U, double* s, double* Vt)
{
    int k = imin(m, n);
    if (m >= n) {
        double* V = dalloc((size_t)n * n);
        svd_tall(m, n, A, U, s, V);
        for (int i = 0; i < k; i++) /* Vt[i][j] = V[j][i]  (col-major Vt k x n) */
            for (int j = 0; j < n; j++) Vt[(long)j * k + i] = V[(long)i * n + j];
        free(V);
    } else {
        /* A^T (n x m) = U' S V'^T  ->  A = V' S U'^T */
        double* At = dalloc((size_t)n * m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < n; j++) At[(long)i * n + j] = A[(long)j * m + i];
        double* Up = dalloc((size_t)n * m);
        double* Vp = dalloc((size_t)m * m);
        svd_tall(n, m, At, Up, s, Vp);
        memcpy(U, Vp, sizeof(double) * m * m);               /* U = V' (m x m) */
        for (int i = 0; i < k; i++)                           /* Vt = U'^T (m x n) */
            for (int j = 0; j < n; j++) Vt[(long)j * k + i] = Up[(long)i * n + j];
        free(At); free(Up); free(Vp);
    }
    return k;
}

/* SVCompressorImp.C:19-73 */
int orc_svcompress(int N, const double* s, int Dmax, double eps, double* discarded,
                   double* rescale)
{
    int D = N;
    double integratedS2 = 0.0;
    if (eps == 0.0) {
        D = Dmax;
        for (int id = D; id < N; id++) integratedS2 += s[id] * s[id];
    } else {
        for (; D >= 1; D--) {
            double s2 = s[D - 1] * s[D - 1];
            if (integratedS2 + s2 >= eps) break;
            integratedS2 += s2;
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
    *discarded = integratedS2;
    *rescale = r;
    return D;
}

/* ------------------------------------------------------------------------ */
/* symmetric tridiagonal: lowest eigenvalues by Sturm bisection and the       */
/* lowest eigenvector by inverse iteration (same algorithm as the device).   */
static int sturm_count(int k, const double* a, const double* b, double x)
{
    /* number of eigenvalues < x */
    int cnt = 0;
    double q = a[0] - x;
    if (q < 0) cnt++;
    for (int i = 1; i < k; i++) {
        if (q == 0.0) q = 1e-300;
        q = a[i] - x - b[i - 1] * b[i - 1] / q;
        if (q < 0) cnt++;
    }
    return cnt;
}

static double tri_eig_bisect(int k, const double* a, const double* b, int idx)
{
    double lo = a[0], hi = a[0];
    for (int i = 0; i < k; i++) {
        double r = (i > 0 ? fabs(b[i - 1]) : 0) + (i < k - 1 ? fabs(b[i]) : 0);
        if (a[i] - r < lo) lo = a[i] - r;
        if (a[i] + r > hi) hi = a[i] + r;
    }
    double span = hi - lo + 1e-300;
    lo -= 1e-10 * span + 1e-300;
    hi += 1e-10 * span + 1e-300;
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_count(k, a, b, mid) > idx) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

/* inverse iteration for eigenvector of T at theta (LU with partial pivoting) */
static void tri_eigvec(int k, const double* a, const double* b, double theta, double* z)
{
    if (k == 1) { z[0] = 1.0; return; }
    double *dl = dalloc(k), *dd = dalloc(k), *du = dalloc(k), *du2 = dalloc(k), *rhs = dalloc(k);
    int* piv = (int*)xcalloc(k, sizeof(int));
    double scale = 0;
    for (int i = 0; i < k; i++) scale = fmax(scale, fabs(a[i]) + (i < k - 1 ? fabs(b[i]) : 0));
    double pert = (scale > 0 ? scale : 1.0) * 1e-14;
    for (int i = 0; i < k; i++) z[i] = 1.0 / sqrt((double)k) * (1.0 + 0.01 * ((i * 7919) % 13) / 13.0);
    for (int iter = 0; iter < 3; iter++) {
        for (int i = 0; i < k; i++) {
            dd[i] = a[i] - theta;
            if (i < k - 1) { du[i] = b[i]; dl[i] = b[i]; }
        }
        /* dgttrf-style LU with partial pivoting */
        for (int i = 0; i < k; i++) du2[i] = 0.0;
        for (int i = 0; i < k - 1; i++) {
            if (fabs(dd[i]) >= fabs(dl[i])) {
                piv[i] = 0;
                if (dd[i] == 0.0) dd[i] = pert;
                double f = dl[i] / dd[i];
                dl[i] = f;
                dd[i + 1] -= f * du[i];
            } else {
                piv[i] = 1;
                double f = dd[i] / dl[i];
                dd[i] = dl[i];
                dl[i] = f;
                double tmp = du[i];
                du[i] = dd[i + 1];
                dd[i + 1] = tmp - f * dd[i + 1];
                if (i < k - 2) {
                    du2[i] = du[i + 1];
                    du[i + 1] = -f * du[i + 1];
                }
            }
        }
        if (dd[k - 1] == 0.0) dd[k - 1] = pert;
        for (int i = 0; i < k; i++) rhs[i] = z[i];
        for (int i = 0; i < k - 1; i++) {
            if (piv[i] == 0) rhs[i + 1] -= dl[i] * rhs[i];
            else { double t = rhs[i]; rhs[i] = rhs[i + 1]; rhs[i + 1] = t - dl[i] * rhs[i + 1]; }
        }
        rhs[k - 1] /= dd[k - 1];
        if (k > 1) rhs[k - 2] = (rhs[k - 2] - du[k - 2] * rhs[k - 1]) / dd[k - 2];
        for (int i = k - 3; i >= 0; i--)
            rhs[i] = (rhs[i] - du[i] * rhs[i + 1] - du2[i] * rhs[i + 2]) / dd[i];
        double nr = 0;
        for (int i = 0; i < k; i++) nr += rhs[i] * rhs[i];
        nr = sqrt(nr);
        for (int i = 0; i < k; i++) z[i] = rhs[i] / nr;
    }
    free(dl); free(dd); free(du); free(du2); free(rhs); free(piv);
}

int orc_lanczos(long N, orc_matvec_fn mv, void* ctx, double* x0, double tol, int maxit,
                double* e0, double* e1)
{
    int kmax = maxit < N ? maxit : (int)N;
    if (kmax < 1) kmax = 1;
    double* V = dalloc((size_t)(kmax + 1) * N);
    double *w = dalloc(N), *alpha = dalloc(kmax + 1), *beta = dalloc(kmax + 1), *z = dalloc(kmax + 1);
    double nr = sqrt(dot(N, x0, x0));
    if (nr == 0.0) {
        uint64_t st = 12345;
        for (long i = 0; i < N; i++) x0[i] = orc_rng_uniform(&st);
        nr = sqrt(dot(N, x0, x0));
    }
    for (long i = 0; i < N; i++) V[i] = x0[i] / nr;
    int nmv = 0, k = 0;
    double theta = 0, theta2 = 0;
    for (int j = 0; j < kmax; j++) {
        double* vj = V + (long)j * N;
        mv(ctx, vj, w);
        nmv++;
        alpha[j] = dot(N, vj, w);
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i <= j; i++) {
                double* vi = V + (long)i * N;
                double c = dot(N, vi, w);
                for (long t = 0; t < N; t++) w[t] -= c * vi[t];
            }
        beta[j] = sqrt(dot(N, w, w));
        k = j + 1;
        theta = tri_eig_bisect(k, alpha, beta, 0);
        theta2 = k > 1 ? tri_eig_bisect(k, alpha, beta, 1) : theta;
        tri_eigvec(k, alpha, beta, theta, z);
        double res = beta[j] * fabs(z[j]);
        if (res < tol || beta[j] < 1e-14 * fmax(1.0, fabs(theta)) || k == N) break;
        double* vn = V + (long)(j + 1) * N;
        for (long t = 0; t < N; t++) vn[t] = w[t] / beta[j];
    }
    for (long t = 0; t < N; t++) {
        double acc = 0;
        for (int i = 0; i < k; i++) acc += z[i] * V[(long)i * N + t];
        x0[t] = acc;
    }
    nr = sqrt(dot(N, x0, x0));
    for (long t = 0; t < N; t++) x0[t] /= nr;
    *e0 = theta;
    if (e1) *e1 = theta2;
    free(V); free(w); free(alpha); free(beta); free(z);
    return nmv;
}

/* ------------------------------------------------------------------------ */
/* exact diagonalisation                                                     */
typedef struct {
    int L, d;
    long N;
    double Jxy, Jz, hz, Dz;
    double *Sp, *Sm, *Sz;
    long* pw;
} ed_ctx;

static void ed_matvec(void* vctx, const double* x, double* y)
{
    ed_ctx* c = (ed_ctx*)vctx;
    const int L = c->L, d = c->d;
    const double S = 0.5 * (d - 1);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < c->N; s++) {
        int dig[64];
        long t = s;
        for (int i = 0; i < L; i++) { dig[i] = (int)(t % d); t /= d; }
        double diag = 0;
        for (int i = 0; i < L; i++) {
            double sz = dig[i] - S;
            diag += c->hz * sz + c->Dz * sz * sz;
            if (i < L - 1) diag += c->Jz * sz * (dig[i + 1] - S);
        }
        double acc = diag * x[s];
        /* y[s] = sum_{s'} H[s][s'] x[s'];  H hermitian: gather from neighbours */
        for (int i = 0; i < L - 1; i++) {
            int a = dig[i], b = dig[i + 1];
            /* <s| S+_i S-_{i+1} |s'>: s' has a-1 at i and b+1 at i+1 */
            if (a >= 1 && b <= d - 2) {
                long sp = s - c->pw[i] + c->pw[i + 1];
                acc += 0.5 * c->Jxy * c->Sp[a * d + (a - 1)] * c->Sm[b * d + (b + 1)] * x[sp];
            }
            if (a <= d - 2 && b >= 1) {
                long sp = s + c->pw[i] - c->pw[i + 1];
                acc += 0.5 * c->Jxy * c->Sm[a * d + (a + 1)] * c->Sp[b * d + (b - 1)] * x[sp];
            }
        }
        y[s] = acc;
    }
}

double orc_ed_ground(int L, int d, double Jxy, double Jz, double hz, double Dz,
                     double tol, int maxit)
{
    ed_ctx c;
    c.L = L; c.d = d; c.Jxy = Jxy; c.Jz = Jz; c.hz = hz; c.Dz = Dz;
    c.N = 1;
    c.pw = (long*)xcalloc(L + 1, sizeof(long));
    for (int i = 0; i < L; i++) { c.pw[i] = c.N; c.N *= d; }
    c.Sp = dalloc(d * d); c.Sm = dalloc(d * d); c.Sz = dalloc(d * d);
    orc_spin_ops(d, c.Sp, c.Sm, c.Sz);
    double* x = dalloc(c.N);
    uint64_t st = 20261202;
    for (long i = 0; i < c.N; i++) x[i] = orc_rng_uniform(&st);
    double e0 = 0, e1 = 0;
    orc_lanczos(c.N, ed_matvec, &c, x, tol, maxit, &e0, &e1);
    free(x); free(c.Sp); free(c.Sm); free(c.Sz); free(c.pw);
    return e0;
}

/* exact diagonalisation of any MPO: y = H x site by site.  The state index
 * has site 0 most significant (s = s_0 d^(L-1) + ... + s_(L-1)).  After site
 * i the work tensor holds X[w][out_0..out_i][in_(i+1)..in_(L-1)]; each step
 * contracts W_i[w][w'][m][n] over (w, n = in_i). */
typedef struct {
    int L, d;
    long N;
    const int *dw1, *dw2;
    const double** W;
    int maxdw;
    double *X, *Y;
} edm_ctx;

static void edm_matvec(void* vctx, const double* x, double* y)
{
    edm_ctx* c = (edm_ctx*)vctx;
    const int L = c->L, d = c->d;
    const long N = c->N;
    memcpy(c->X, x, sizeof(double) * N); /* w = 0 only (dw1[0] = 1) */
    long head = 1;                       /* d^i: processed (output) sites */
    for (int i = 0; i < L; i++) {
        const int a = c->dw1[i], b = c->dw2[i];
        const long tail = N / (head * d); /* remaining input sites after i */
        const double* W = c->W[i];
        memset(c->Y, 0, sizeof(double) * (size_t)b * N);
        /* X[w][h][n][t] -> Y[w'][h][m][t] */
#pragma omp parallel for schedule(static)
        for (long h = 0; h < head; h++)
            for (int w = 0; w < a; w++)
                for (int w2 = 0; w2 < b; w2++)
                    for (int m = 0; m < d; m++)
                        for (int n = 0; n < d; n++) {
                            const double v = W[((w * b + w2) * d + m) * d + n];
                            if (v == 0.0) continue;
                            const double* src = c->X + (size_t)w * N + (h * d + n) * tail;
                            double* dst = c->Y + (size_t)w2 * N + (h * d + m) * tail;
                            for (long t = 0; t < tail; t++) dst[t] += v * src[t];
                        }
        double* tmp = c->X;
        c->X = c->Y;
        c->Y = tmp;
        head *= d;
    }
    memcpy(y, c->X, sizeof(double) * N); /* dw2[L-1] = 1 */
}

double orc_ed_ground_mpo(int L, int d, const int* dw1, const int* dw2, const double* wcat, double tol,
                         int maxit)
{
    edm_ctx c;
    c.L = L; c.d = d; c.dw1 = dw1; c.dw2 = dw2;
    c.N = 1;
    for (int i = 0; i < L; i++) c.N *= d;
    c.W = (const double**)xcalloc(L, sizeof(double*));
    c.maxdw = 1;
    size_t off = 0;
    for (int i = 0; i < L; i++) {
        c.W[i] = wcat + off;
        off += (size_t)dw1[i] * dw2[i] * d * d;
        if (dw1[i] > c.maxdw) c.maxdw = dw1[i];
        if (dw2[i] > c.maxdw) c.maxdw = dw2[i];
    }
    c.X = dalloc((size_t)c.maxdw * c.N);
    c.Y = dalloc((size_t)c.maxdw * c.N);
    double* x = dalloc(c.N);
    uint64_t st = 20261202;
    for (long i = 0; i < c.N; i++) x[i] = orc_rng_uniform(&st);
    double e0 = 0, e1 = 0;
    orc_lanczos(c.N, edm_matvec, &c, x, tol, maxit, &e0, &e1);
    free(x); free(c.X); free(c.Y); free((void*)c.W);
    return e0;
}

/* ------------------------------------------------------------------------ */
/* MPS + sweep: MPSImp.C, MPSSite.C, Bond.C, MPSImp_Variational.C            */
struct orc_mps {
    int L, d, form;
    int *D1, *D2;
    double** M;          /* reference layout per site */
    int *Dw1, *Dw2;
    double** W;          /* per site H operator */
    int *Ew1, *Ew2;
    double** W2;         /* per site H^2 operator (uncompressed product) */
    double** Lc;         /* left cache of site i: Dw2(i) x D2(i) x D2(i) */
    double** Rc;         /* right cache of site i: Dw1(i) x D1(i) x D1(i) */
    double *Emin, *IterDE, *GapE;
    /* bonds b=0..L-2 */
    int* bD;
    double** lam;
    double* maxDelta;
    double* compErr;
    int nmatvec;
    double lastE;
};

static void site_resize(orc_mps* s, int i, int D1, int D2, int save)
{
    /* MPSSite.C:156-171 NewBondDimensions: keep the overlapping block when
     * saving, then zero entries (i1>=oldD1-1, i2>=oldD2-1) like the reference's
     * fill loop (1-based from the old extent). */
    const int d = s->d;
    int oD1 = s->D1[i], oD2 = s->D2[i];
    if (oD1 == D1 && oD2 == D2) return;
    double* nM = dalloc((size_t)d * D1 * D2);
    if (save) {
        for (int m = 0; m < d; m++)
            for (int i1 = 0; i1 < imin(oD1, D1); i1++)
                for (int i2 = 0; i2 < imin(oD2, D2); i2++)
                    nM[m + d * (i1 + D1 * i2)] = s->M[i][m + d * (i1 + oD1 * i2)];
        for (int m = 0; m < d; m++)
            for (int i1 = oD1 - 1; i1 < D1; i1++)
                for (int i2 = oD2 - 1; i2 < D2; i2++)
                    if (i1 >= 0 && i2 >= 0) nM[m + d * (i1 + D1 * i2)] = 0.0;
    }
    free(s->M[i]);
    s->M[i] = nM;
    s->D1[i] = D1;
    s->D2[i] = D2;
}

static void site_random(orc_mps* s, int i, uint64_t* st)
{
    const int d = s->d, D1 = s->D1[i], D2 = s->D2[i];
    double sc = 1.0 / sqrt((double)d * D1 * D2);
    for (int m = 0; m < d; m++)
        for (int i2 = 0; i2 < D2; i2++)
            for (int i1 = 0; i1 < D1; i1++) s->M[i][PSI(m, i1, i2)] = orc_rng_uniform(st) * sc;
}

orc_mps* orc_mps_create(int L, int d, double Jxy, double Jz, double hz, double Dz,
                        int form, int D, int init, uint64_t seed)
{
    orc_mps* s = (orc_mps*)xcalloc(1, sizeof(orc_mps));
    s->L = L; s->d = d; s->form = form;
    s->D1 = (int*)xcalloc(L, sizeof(int)); s->D2 = (int*)xcalloc(L, sizeof(int));
    s->M = (double**)xcalloc(L, sizeof(double*));
    s->Dw1 = (int*)xcalloc(L, sizeof(int)); s->Dw2 = (int*)xcalloc(L, sizeof(int));
    s->Ew1 = (int*)xcalloc(L, sizeof(int)); s->Ew2 = (int*)xcalloc(L, sizeof(int));
    s->W = (double**)xcalloc(L, sizeof(double*)); s->W2 = (double**)xcalloc(L, sizeof(double*));
    s->Lc = (double**)xcalloc(L, sizeof(double*)); s->Rc = (double**)xcalloc(L, sizeof(double*));
    s->Emin = dalloc(L); s->IterDE = dalloc(L); s->GapE = dalloc(L);
    int nb = L > 1 ? L - 1 : 1;
    s->bD = (int*)xcalloc(nb, sizeof(int));
    s->lam = (double**)xcalloc(nb, sizeof(double*));
    s->maxDelta = dalloc(nb); s->compErr = dalloc(nb);
    double* Wb = dalloc(25 * d * d);
    orc_xxz_bulk_W(d, Jxy, Jz, hz, Dz, form, Wb);
    for (int i = 0; i < L; i++) {
        int pos = (L == 1) ? 3 : (i == 0 ? 0 : (i == L - 1 ? 2 : 1));
        s->W[i] = dalloc(25 * d * d);
        orc_site_W(d, 5, Wb, form, pos, s->W[i], &s->Dw1[i], &s->Dw2[i]);
        s->Ew1[i] = s->Dw1[i] * s->Dw1[i];
        s->Ew2[i] = s->Dw2[i] * s->Dw2[i];
        s->W2[i] = dalloc((size_t)s->Ew1[i] * s->Ew2[i] * d * d);
        orc_mpo_product(d, s->Dw1[i], s->Dw2[i], s->W[i], s->Dw1[i], s->Dw2[i], s->W[i],
                        s->W2[i]);
        /* MPSImp.C:69-90: site1 1xD, bulk DxD, siteL Dx1 */
        s->D1[i] = (i == 0) ? 1 : D;
        s->D2[i] = (i == L - 1) ? 1 : D;
        s->M[i] = dalloc((size_t)d * s->D1[i] * s->D2[i]);
        s->IterDE[i] = 1.0; /* MPSSite.C:40 */
    }
    free(Wb);
    for (int b = 0; b < L - 1; b++) { /* Bond.C:11-25 */
        s->bD[b] = D;
        s->lam[b] = dalloc(D);
        for (int k = 0; k < D; k++) s->lam[b][k] = 1.0 / D;
        s->compErr[b] = -99;
    }
    uint64_t st = seed;
    int sgn = 1;
    for (int i = 0; i < L; i++) { /* MPSSite.C:84-140 */
        int D1 = s->D1[i], D2 = s->D2[i];
        double* M = s->M[i];
        if (init == 1) site_random(s, i, &st);
        else if (init == 2) {
            if (sgn == 1) M[PSI(0, 0, 0)] = 1.0;
            else M[PSI(d - 1, 0, 0)] = 1.0;
        } else {
            if (i == 0 && L > 1) { for (int m = 0; m < d && m < D2; m++) M[PSI(m, 0, m)] = sgn; }
            else if (i == L - 1 && L > 1) { for (int m = 0; m < d && m < D1; m++) M[PSI(m, m, 0)] = sgn; }
            else for (int m = 0; m < d; m++) for (int k = 0; k < imin(D1, D2); k++) M[PSI(m, k, k)] = sgn / sqrt((double)d);
        }
        sgn = -sgn;
    }
    return s;
}

void orc_mps_destroy(orc_mps* s)
{
    if (!s) return;
    for (int i = 0; i < s->L; i++) {
        free(s->M[i]); free(s->W[i]); free(s->W2[i]); free(s->Lc[i]); free(s->Rc[i]);
    }
    for (int b = 0; b < s->L - 1; b++) free(s->lam[b]);
    free(s->D1); free(s->D2); free(s->M); free(s->Dw1); free(s->Dw2); free(s->Ew1);
    free(s->Ew2); free(s->W); free(s->W2); free(s->Lc); free(s->Rc); free(s->Emin);
    free(s->IterDE); free(s->GapE); free(s->bD); free(s->lam); free(s->maxDelta);
    free(s->compErr);
    free(s);
}

static double norm_vec(long n, const double* v) { return sqrt(dot(n, v, v)); }

/* Bond.C:86-106 SetSingularValues */
static void bond_set_sv(orc_mps* s, int b, const double* sv, int n, double comp)
{
    double nr = norm_vec(n, sv);
    double* sn = dalloc(n);
    for (int i = 0; i < n; i++) sn[i] = sv[i] / nr;
    double* old = dalloc(n);
    for (int i = 0; i < n && i < s->bD[b]; i++) old[i] = s->lam[b][i];
    double md = 0;
    for (int i = 0; i < n; i++) md = fmax(md, fabs(sn[i] - old[i]));
    s->maxDelta[b] = md;
    free(s->lam[b]);
    free(old);
    s->lam[b] = sn;
    s->bD[b] = n;
    s->compErr[b] += comp;
}

/* Bond.C:60-84 NewBondDimension */
static void bond_new_dim(orc_mps* s, int b, int D)
{
    int oD = s->bD[b];
    double* nl = dalloc(D);
    if (D > oD) {
        double mn = s->lam[b][0];
        for (int i = 1; i < oD; i++) mn = fmin(mn, s->lam[b][i]);
        for (int i = 0; i < oD; i++) nl[i] = s->lam[b][i];
        for (int i = oD; i < D; i++) nl[i] = mn / 10.0;
    } else {
        for (int i = 0; i < D; i++) nl[i] = s->lam[b][i];
    }
    double nr = norm_vec(D, nl);
    for (int i = 0; i < D; i++) nl[i] /= nr;
    free(s->lam[b]);
    s->lam[b] = nl;
    s->bD[b] = D;
}

static double site_norm2(orc_mps* s, int i)
{
    return dot((long)s->d * s->D1[i] * s->D2[i], s->M[i], s->M[i]);
}

static void site_rescale(orc_mps* s, int i, double f)
{
    long n = (long)s->d * s->D1[i] * s->D2[i];
    for (long k = 0; k < n; k++) s->M[i][k] /= f;
}

/* MPSSite_Normalization.C:20-78 SVDNormalize (compressor = nullptr) */
static void svd_normalize(orc_mps* s, int i, int dir)
{
    const int d = s->d, L = s->L;
    int D1 = s->D1[i], D2 = s->D2[i];
    if (dir == 1 && i == 0) { /* HandleEdges DRight */
        int newD2 = (L > 1) ? imax(s->bD[0], d) : D2;
        if (newD2 < D2) site_resize(s, i, D1, newD2, 1);
        site_rescale(s, i, sqrt(site_norm2(s, i)));
        return;
    }
    if (dir == 0 && i == L - 1) { /* HandleEdges DLeft */
        int newD1 = (L > 1) ? imax(s->bD[L - 2], d) : D1;
        if (newD1 < D1) site_resize(s, i, newD1, D2, 1);
        site_rescale(s, i, sqrt(site_norm2(s, i)));
        return;
    }
    if (dir == 0) {
        /* ReshapeBeforeSVD DLeft: A(n*D1+i1, i2) (col-major, rows d*D1) */
        int m = d * D1, n = D2, k = imin(m, n);
        double* A = dalloc((size_t)m * n);
        for (int nn = 0; nn < d; nn++)
            for (int i1 = 0; i1 < D1; i1++)
                for (int i2 = 0; i2 < D2; i2++)
                    A[(long)i2 * m + nn * D1 + i1] = s->M[i][PSI(nn, i1, i2)];
        double *U = dalloc((size_t)m * k), *sv = dalloc(k), *Vt = dalloc((size_t)k * n);
        orc_svd(m, n, A, U, sv, Vt);
        /* Bond::SVDTransfer -> SetSingularValues; neighbour MPSSite::SVDTransfer DLeft */
        bond_set_sv(s, i, sv, k, 0.0);
        int j = i + 1;
        int jD1 = s->D1[j], jD2 = s->D2[j];
        double* nM = dalloc((size_t)d * k * jD2);
        for (int nn = 0; nn < d; nn++)
            for (int a = 0; a < k; a++)
                for (int c = 0; c < jD2; c++) {
                    double acc = 0;
                    for (int b = 0; b < jD1; b++)
                        acc += Vt[(long)b * k + a] * s->M[j][nn + d * (b + jD1 * c)];
                    nM[nn + d * (a + k * c)] = sv[a] * acc;
                }
        free(s->M[j]);
        s->M[j] = nM;
        s->D1[j] = k;
        /* ReshapeAfter_SVD DLeft */
        double* sM = dalloc((size_t)d * D1 * k);
        for (int nn = 0; nn < d; nn++)
            for (int i1 = 0; i1 < D1; i1++)
                for (int i2 = 0; i2 < k; i2++)
                    sM[nn + d * (i1 + D1 * i2)] = U[(long)i2 * m + nn * D1 + i1];
        free(s->M[i]);
        s->M[i] = sM;
        s->D2[i] = k;
        free(A); free(U); free(sv); free(Vt);
    } else {
        /* ReshapeBeforeSVD DRight: A(i1, n*D2+i2) rows D1 */
        int m = D1, n = d * D2, k = imin(m, n);
        double* A = dalloc((size_t)m * n);
        for (int nn = 0; nn < d; nn++)
            for (int i2 = 0; i2 < D2; i2++)
                for (int i1 = 0; i1 < D1; i1++)
                    A[(long)(nn * D2 + i2) * m + i1] = s->M[i][PSI(nn, i1, i2)];
        double *U = dalloc((size_t)m * k), *sv = dalloc(k), *Vt = dalloc((size_t)k * n);
        orc_svd(m, n, A, U, sv, Vt);
        bond_set_sv(s, i - 1, sv, k, 0.0);
        int j = i - 1;
        int jD1 = s->D1[j], jD2 = s->D2[j];
        double* nM = dalloc((size_t)d * jD1 * k);
        for (int nn = 0; nn < d; nn++)
            for (int a = 0; a < jD1; a++)
                for (int c = 0; c < k; c++) {
                    double acc = 0;
                    for (int b = 0; b < jD2; b++)
                        acc += s->M[j][nn + d * (a + jD1 * b)] * U[(long)c * m + b];
                    nM[nn + d * (a + jD1 * c)] = acc * sv[c];
                }
        free(s->M[j]);
        s->M[j] = nM;
        s->D2[j] = k;
        double* sM = dalloc((size_t)d * k * D2);
        for (int nn = 0; nn < d; nn++)
            for (int i2 = 0; i2 < D2; i2++)
                for (int i1 = 0; i1 < k; i1++)
                    sM[nn + d * (i1 + k * i2)] = Vt[(long)(nn * D2 + i2) * k + i1];
        free(s->M[i]);
        s->M[i] = sM;
        s->D1[i] = k;
        free(A); free(U); free(sv); free(Vt);
    }
}

void orc_mps_normalize(orc_mps* s, int dir)
{
    if (dir == 0) for (int i = 0; i < s->L; i++) svd_normalize(s, i, 0);
    else for (int i = s->L - 1; i >= 0; i--) svd_normalize(s, i, 1);
}

static const double ONE = 1.0;

static void env_left_of(orc_mps* s, int i, const double* Fin, double* Fout, int which)
{
    const double* W = which == 0 ? s->W[i] : s->W2[i];
    int Dw1 = which == 0 ? s->Dw1[i] : s->Ew1[i], Dw2 = which == 0 ? s->Dw2[i] : s->Ew2[i];
    orc_env_left_ref(s->d, s->D1[i], s->D2[i], Dw1, Dw2, W, Fin, s->M[i], Fout);
}

/* fast left/right env of site i via the GEMM-ordered path (same sums) */
static void env_update_fast(orc_mps* s, int i, int dir)
{
    const int d = s->d, D1 = s->D1[i], D2 = s->D2[i], Dw1 = s->Dw1[i], Dw2 = s->Dw2[i];
    double* A = dalloc((size_t)d * D1 * D2);
    orc_psi_ref_to_int(d, D1, D2, s->M[i], A);
    double* work = dalloc(orc_heff_work_size(d, D1, D2, Dw1, Dw2));
    if (dir == 0) {
        const double* Fin = (i > 0) ? s->Lc[i - 1] : &ONE;
        double* Lin = dalloc((size_t)Dw1 * D1 * D1);
        orc_lenv_ref_to_int(Dw1, D1, Fin, Lin);
        double* Lo = dalloc((size_t)Dw2 * D2 * D2);
        orc_env_left_int(d, D1, D2, Dw1, Dw2, s->W[i], Lin, A, Lo, work);
        free(s->Lc[i]);
        s->Lc[i] = dalloc((size_t)Dw2 * D2 * D2);
        orc_lenv_int_to_ref(Dw2, D2, Lo, s->Lc[i]);
        free(Lin); free(Lo);
    } else {
        const double* Fin = (i < s->L - 1) ? s->Rc[i + 1] : &ONE;
        double* Rin = dalloc((size_t)Dw2 * D2 * D2);
        orc_renv_ref_to_int(Dw2, D2, Fin, Rin);
        double* Ro = dalloc((size_t)Dw1 * D1 * D1);
        orc_env_right_int(d, D1, D2, Dw1, Dw2, s->W[i], Rin, A, Ro, work);
        free(s->Rc[i]);
        s->Rc[i] = dalloc((size_t)Dw1 * D1 * D1);
        orc_renv_int_to_ref(Dw1, D1, Ro, s->Rc[i]);
        free(Rin); free(Ro);
    }
    free(A); free(work);
}

static void update_cache(orc_mps* s, int i, int mode)
{
    /* MPSSite_Variational.C:46-50: both L and R caches of the visited site */
    if (mode == 0) {
        const double* Lin = (i > 0) ? s->Lc[i - 1] : &ONE;
        const double* Rin = (i < s->L - 1) ? s->Rc[i + 1] : &ONE;
        free(s->Lc[i]);
        s->Lc[i] = dalloc((size_t)s->Dw2[i] * s->D2[i] * s->D2[i]);
        orc_env_left_ref(s->d, s->D1[i], s->D2[i], s->Dw1[i], s->Dw2[i], s->W[i], Lin,
                         s->M[i], s->Lc[i]);
        free(s->Rc[i]);
        s->Rc[i] = dalloc((size_t)s->Dw1[i] * s->D1[i] * s->D1[i]);
        orc_env_right_ref(s->d, s->D1[i], s->D2[i], s->Dw1[i], s->Dw2[i], s->W[i], Rin,
                          s->M[i], s->Rc[i]);
    } else {
        env_update_fast(s, i, 0);
        env_update_fast(s, i, 1);
    }
}

/* MPSImp_Variational.C:120-149 LoadHeffCaches */
void orc_mps_load_caches(orc_mps* s)
{
    for (int i = 0; i < s->L - 1; i++) env_update_fast(s, i, 0);
    for (int i = s->L - 1; i >= 1; i--) env_update_fast(s, i, 1);
}

typedef struct {
    int d, D1, D2, Dw1, Dw2;
    const double *W, *L, *R;
    double* work;
    double *xi, *yi;
} heff_ctx;

static void heff_mv(void* vc, const double* x, double* y)
{
    heff_ctx* c = (heff_ctx*)vc;
    orc_psi_ref_to_int(c->d, c->D1, c->D2, x, c->xi);
    orc_heff_apply_int(c->d, c->D1, c->D2, c->Dw1, c->Dw2, c->W, c->L, c->R, c->xi, c->yi,
                       c->work);
    orc_psi_int_to_ref(c->d, c->D1, c->D2, c->yi, y);
}

/* MPSImp_Variational.C:79-99 Refine + MPSSite_Variational.C:18-44 */
static void refine(orc_mps* s, int dir, int i, double epsEig, int mode)
{
    const int d = s->d, D1 = s->D1[i], D2 = s->D2[i], Dw1 = s->Dw1[i], Dw2 = s->Dw2[i];
    const long N = (long)d * D1 * D2;
    const double* Lr = (i > 0) ? s->Lc[i - 1] : &ONE;
    const double* Rr = (i < s->L - 1) ? s->Rc[i + 1] : &ONE;
    double e0, e1;
    if (mode == 0) {
        double* H = dalloc((size_t)N * N);
        orc_heff_dense(d, D1, D2, Dw1, Dw2, s->W[i], Lr, Rr, H);
        double *w = dalloc(N), *Z = dalloc((size_t)N * N);
        orc_syev((int)N, H, w, Z);
        e0 = w[0];
        e1 = N > 1 ? w[1] : w[0];
        memcpy(s->M[i], Z, sizeof(double) * N);
        free(H); free(w); free(Z);
        s->nmatvec += 1;
    } else {
        heff_ctx c;
        c.d = d; c.D1 = D1; c.D2 = D2; c.Dw1 = Dw1; c.Dw2 = Dw2; c.W = s->W[i];
        double* Li = dalloc((size_t)Dw1 * D1 * D1);
        double* Ri = dalloc((size_t)Dw2 * D2 * D2);
        orc_lenv_ref_to_int(Dw1, D1, Lr, Li);
        orc_renv_ref_to_int(Dw2, D2, Rr, Ri);
        c.L = Li; c.R = Ri;
        c.work = dalloc(orc_heff_work_size(d, D1, D2, Dw1, Dw2));
        c.xi = dalloc(N); c.yi = dalloc(N);
        s->nmatvec += orc_lanczos(N, heff_mv, &c, s->M[i], epsEig, 200, &e0, &e1);
        free(Li); free(Ri); free(c.work); free(c.xi); free(c.yi);
    }
    s->IterDE[i] = e0 - s->Emin[i];
    s->Emin[i] = e0;
    s->GapE[i] = e1 - 1.0; /* MPSSite_Variational.C:29 as written */
    s->lastE = e0;
    svd_normalize(s, i, dir);
    update_cache(s, i, mode);
}

void orc_mps_sweep(orc_mps* s, int dir, double epsEig, int mode)
{
    if (dir == 0) for (int i = 0; i < s->L; i++) refine(s, 0, i, epsEig, mode);
    else for (int i = s->L - 1; i >= 0; i--) refine(s, 1, i, epsEig, mode);
}

double orc_mps_max_delta_e(const orc_mps* s)
{
    double m = 0;
    for (int i = 0; i < s->L; i++) m = fmax(m, fabs(s->IterDE[i]));
    return m;
}
double orc_mps_max_delta_l(const orc_mps* s)
{
    double m = 0;
    for (int b = 0; b < s->L - 1; b++) m = fmax(m, s->maxDelta[b]);
    return m;
}

/* MPSImp.C:123-147 IncreaseBondDimensions */
static void increase_bond_dims(orc_mps* s, int D)
{
    const int L = s->L, d = s->d;
    long Dideal = 1;
    int ia = 0;
    for (; ia < L / 2; ia++) {
        int D1 = (int)(Dideal < D ? Dideal : D), D2 = (int)(Dideal * d < D ? Dideal * d : D);
        site_resize(s, ia, D1, D2, 1);
        bond_new_dim(s, ia, D2);
        Dideal *= d;
    }
    Dideal = 1;
    for (int ib = L - 1; ib >= ia; ib--) {
        int D2 = (int)(Dideal < D ? Dideal : D), D1 = (int)(Dideal * d < D ? Dideal * d : D);
        site_resize(s, ib, D1, D2, 1);
        bond_new_dim(s, ib - 1, D1);
        Dideal *= d;
    }
}

double orc_mps_expectation(orc_mps* s, int which)
{
    /* MPSImp_ExpectationsAndDMs.C:33-45 */
    double* F = dalloc(1);
    F[0] = 1.0;
    for (int i = 0; i < s->L; i++) {
        int Dw2 = which == 1 ? s->Ew2[i] : s->Dw2[i];
        double* Fn = dalloc((size_t)Dw2 * s->D2[i] * s->D2[i]);
        if (which == 2) {
            /* identity MPO: Dw=1, W = I */
            double* Wi = dalloc(s->d * s->d);
            for (int m = 0; m < s->d; m++) Wi[m * s->d + m] = 1.0;
            double* Fi = dalloc((size_t)s->D2[i] * s->D2[i]);
            orc_env_left_ref(s->d, s->D1[i], s->D2[i], 1, 1, Wi, F, s->M[i], Fi);
            free(Fn);
            Fn = Fi;
            free(Wi);
        } else {
            env_left_of(s, i, F, Fn, which);
        }
        free(F);
        F = Fn;
    }
    double r = F[0];
    free(F);
    return r;
}

static int max_D(const orc_mps* s)
{
    int D = 0;
    for (int i = 0; i < s->L; i++) D = imax(D, s->D2[i]);
    return D;
}

/* MPSImp_Variational.C:27-62 */
double orc_mps_find_gs_line(orc_mps* s, int maxiter, int Dmax, int DeltaD, double epsE,
                            double epsL, double epsEig, int mode, int* nsweeps)
{
    int D0 = max_D(s);
    double DE2 = 0;
    int in = 0;
    for (int D = D0; D <= Dmax; D += DeltaD) {
        if (D > D0) {
            increase_bond_dims(s, D);
            orc_mps_normalize(s, 1);
            orc_mps_load_caches(s);
        }
        for (in = 0; in < maxiter; in++) {
            orc_mps_sweep(s, 0, epsEig, mode);
            orc_mps_sweep(s, 1, epsEig, mode);
            double dE = orc_mps_max_delta_e(s), dl = orc_mps_max_delta_l(s);
            if (dE < epsE && dl < epsL) break;
        }
        in++;
        double E1 = orc_mps_expectation(s, 0);
        double E2 = orc_mps_expectation(s, 1);
        DE2 = E2 - E1 * E1;
    }
    if (nsweeps) *nsweeps = in;
    return DE2;
}

int orc_mps_get_dims(const orc_mps* s, int* D1, int* D2)
{
    for (int i = 0; i < s->L; i++) { D1[i] = s->D1[i]; D2[i] = s->D2[i]; }
    return s->L;
}
int orc_mps_get_site(const orc_mps* s, int i, double* M)
{
    long n = (long)s->d * s->D1[i] * s->D2[i];
    memcpy(M, s->M[i], sizeof(double) * n);
    return (int)n;
}
int orc_mps_set_site(orc_mps* s, int i, int D1, int D2, const double* M)
{
    free(s->M[i]);
    s->D1[i] = D1; s->D2[i] = D2;
    long n = (long)s->d * D1 * D2;
    s->M[i] = dalloc(n);
    memcpy(s->M[i], M, sizeof(double) * n);
    return 0;
}
int orc_mps_total_matvecs(const orc_mps* s) { return s->nmatvec; }
double orc_mps_last_energy(const orc_mps* s) { return s->lastE; }
int orc_mps_get_bond_sv(const orc_mps* s, int b, double* sv)
{
    memcpy(sv, s->lam[b], sizeof(double) * s->bD[b]);
    return s->bD[b];
}

/* MPSSite.C:173-200 GetNormStatus: 'A' left-, 'B' right-normalised, 'I' both,
 * 'M' neither (eps 1e-12) */
void orc_mps_norm_status(const orc_mps* s, char* out)
{
    const int d = s->d;
    for (int i = 0; i < s->L; i++) {
        int D1 = s->D1[i], D2 = s->D2[i];
        const double* M = s->M[i];
        double errA = 0, errB = 0;
        for (int a = 0; a < D2; a++)
            for (int b = 0; b < D2; b++) {
                double acc = 0;
                for (int m = 0; m < d; m++)
                    for (int i1 = 0; i1 < D1; i1++) acc += M[PSI(m, i1, a)] * M[PSI(m, i1, b)];
                errA = fmax(errA, fabs(acc - (a == b)));
            }
        for (int a = 0; a < D1; a++)
            for (int b = 0; b < D1; b++) {
                double acc = 0;
                for (int m = 0; m < d; m++)
                    for (int i2 = 0; i2 < D2; i2++) acc += M[PSI(m, a, i2)] * M[PSI(m, b, i2)];
                errB = fmax(errB, fabs(acc - (a == b)));
            }
        int A = errA < 1e-12, B = errB < 1e-12;
        out[i] = (A && B) ? 'I' : (A ? 'A' : (B ? 'B' : 'M'));
    }
    out[s->L] = 0;
}

/* ------------------------------------------------------------------------ */
/* two-site sweep (extension; internal layout Theta[a][m1][m2][c])           */
static void two_site_W(int d, int Dw1, int Dwm, int Dw3, const double* Wa, const double* Wb,
                       double* W12)
{
    /* W12[w][w''][(m1,m2)][(n1,n2)] = sum_w' Wa[w][w'][m1][n1] Wb[w'][w''][m2][n2] */
    const int p = d * d;
    memset(W12, 0, sizeof(double) * Dw1 * Dw3 * p * p);
    for (int w = 0; w < Dw1; w++)
        for (int w2 = 0; w2 < Dw3; w2++)
            for (int wp = 0; wp < Dwm; wp++)
                for (int m1 = 0; m1 < d; m1++)
                    for (int n1 = 0; n1 < d; n1++) {
                        double a = Wa[((w * Dwm + wp) * d + m1) * d + n1];
                        if (a == 0.0) continue;
                        for (int m2 = 0; m2 < d; m2++)
                            for (int n2 = 0; n2 < d; n2++)
                                W12[((w * Dw3 + w2) * p + (m1 * d + m2)) * p + (n1 * d + n2)] +=
                                    a * Wb[((wp * Dw3 + w2) * d + m2) * d + n2];
                    }
}

double orc_mps_sweep2(orc_mps* s, int dir, int Dmax, double epsSV, double epsEig)
{
    const int d = s->d, L = s->L, p = d * d;
    double maxTrunc = 0;
    for (int step = 0; step < L - 1; step++) {
        int i = (dir == 0) ? step : L - 2 - step;
        int j = i + 1;
        int D1 = s->D1[i], Dm = s->D2[i], D3 = s->D2[j];
        int Dw1 = s->Dw1[i], Dwm = s->Dw2[i], Dw3 = s->Dw2[j];
        double* Ai = dalloc((size_t)D1 * d * Dm);
        double* Aj = dalloc((size_t)Dm * d * D3);
        orc_psi_ref_to_int(d, D1, Dm, s->M[i], Ai);
        orc_psi_ref_to_int(d, Dm, D3, s->M[j], Aj);
        long N = (long)D1 * p * D3;
        double* th = dalloc(N); /* [a][m1][m2][c] = Ai[(a,m1)][b] Aj[b][(m2,c)] */
        orc_gemm(D1 * d, d * D3, Dm, Ai, Dm, 1, Aj, (long)d * D3, 1, 0.0, th, (long)d * D3);
        double* W12 = dalloc((size_t)Dw1 * Dw3 * p * p);
        two_site_W(d, Dw1, Dwm, Dw3, s->W[i], s->W[j], W12);
        const double* Lr = (i > 0) ? s->Lc[i - 1] : &ONE;
        const double* Rr = (j < L - 1) ? s->Rc[j + 1] : &ONE;
        double* Li = dalloc((size_t)Dw1 * D1 * D1);
        double* Ri = dalloc((size_t)Dw3 * D3 * D3);
        orc_lenv_ref_to_int(Dw1, D1, Lr, Li);
        orc_renv_ref_to_int(Dw3, D3, Rr, Ri);
        /* Lanczos over internal layout directly (vector = th) */
        heff_ctx c;
        c.d = p; c.D1 = D1; c.D2 = D3; c.Dw1 = Dw1; c.Dw2 = Dw3; c.W = W12; c.L = Li; c.R = Ri;
        c.work = dalloc(orc_heff_work_size(p, D1, D3, Dw1, Dw3));
        c.xi = dalloc(N); c.yi = dalloc(N);
        /* heff_mv converts ref<->int; for the two-site vector use the int layout
         * on both sides by passing ref=int through the converters' inverse pair */
        double* thr = dalloc(N);
        orc_psi_int_to_ref(p, D1, D3, th, thr);
        double e0, e1;
        s->nmatvec += orc_lanczos(N, heff_mv, &c, thr, epsEig, 200, &e0, &e1);
        orc_psi_ref_to_int(p, D1, D3, thr, th);
        s->lastE = e0;
        /* SVD of Theta as (D1*d) x (d*D3), row-major == col-major of transpose */
        int m = D1 * d, n = d * D3, k = imin(m, n);
        double* A = dalloc((size_t)m * n);
        for (int r = 0; r < m; r++)
            for (int q = 0; q < n; q++) A[(long)q * m + r] = th[(long)r * n + q];
        double *U = dalloc((size_t)m * k), *sv = dalloc(k), *Vt = dalloc((size_t)k * n);
        orc_svd(m, n, A, U, sv, Vt);
        double disc = 0, resc = 1;
        int Dk = orc_svcompress(k, sv, Dmax, epsSV, &disc, &resc);
        if (Dk < 1) Dk = 1;
        if (Dk < k) for (int t = 0; t < Dk; t++) sv[t] *= resc;
        if (disc > maxTrunc) maxTrunc = disc;
        bond_set_sv(s, i, sv, Dk, disc);
        double* Ni = dalloc((size_t)D1 * d * Dk);
        double* Nj = dalloc((size_t)Dk * d * D3);
        for (int r = 0; r < m; r++)
            for (int t = 0; t < Dk; t++)
                Ni[(long)r * Dk + t] = U[(long)t * m + r] * (dir == 1 ? sv[t] : 1.0);
        double nrm = 0;
        for (int t = 0; t < Dk; t++) nrm += sv[t] * sv[t];
        nrm = sqrt(nrm);
        for (int t = 0; t < Dk; t++)
            for (int q = 0; q < n; q++)
                Nj[(long)t * n + q] = Vt[(long)q * k + t] * (dir == 0 ? sv[t] : 1.0);
        /* keep the state normalised after truncation */
        if (dir == 0) for (long t = 0; t < (long)Dk * n; t++) Nj[t] /= nrm;
        else for (long t = 0; t < (long)m * Dk; t++) Ni[t] /= nrm;
        free(s->M[i]); free(s->M[j]);
        s->M[i] = dalloc((size_t)D1 * d * Dk);
        s->M[j] = dalloc((size_t)Dk * d * D3);
        s->D2[i] = Dk; s->D1[j] = Dk;
        orc_psi_int_to_ref(d, D1, Dk, Ni, s->M[i]);
        orc_psi_int_to_ref(d, Dk, D3, Nj, s->M[j]);
        if (dir == 0) env_update_fast(s, i, 0);
        else env_update_fast(s, j, 1);
        free(Ai); free(Aj); free(th); free(W12); free(Li); free(Ri); free(c.work);
        free(c.xi); free(c.yi); free(thr); free(A); free(U); free(sv); free(Vt);
        free(Ni); free(Nj);
    }
    return maxTrunc;
}


/* ---- learned Davidson preconditioner: the network's definition ---------- */
void orc_mlp_features(long n, const double* diag, double theta, const double* lam_l, const double* lam_r,
                      double delta, double* X)
{
    /* theta enters relative to the spread of the diagonal (an intensive
     * quantity; theta itself grows with the chain length) */
    double mu = 0.0, sq = 0.0;
    for (long i = 0; i < n; i++) {
        mu += diag[i];
        sq += diag[i] * diag[i];
    }
    mu = n > 0 ? mu / (double)n : 0.0;
    double sd = n > 0 ? sqrt(fmax(sq / (double)n - mu * mu, 0.0)) : 0.0;
    if (!(sd > 1e-12)) sd = 1.0;
    for (long i = 0; i < n; i++) {
        X[4 * i + 0] = asinh((diag[i] - theta) / delta);
        X[4 * i + 1] = (theta - mu) / sd;
        X[4 * i + 2] = log(lam_l[i] + 1e-10);
        X[4 * i + 3] = log(lam_r[i] + 1e-10);
    }
}

/* one sample: forward, optionally backward (dy = dloss/dy, grad accumulated) */
static double mlp_one(const double* P, const double* norm, const double* x, double dy_scale_a, double t,
                      double* grad, double* g_out)
{
    const int H = ORC_MLP_H;
    const double* W1 = P;
    const double* b1 = W1 + H * 4;
    const double* W2 = b1 + H;
    const double* b2 = W2 + H * H;
    const double* w3 = b2 + H;
    const double b3 = w3[H];
    double xh[4], h1[ORC_MLP_H], h2[ORC_MLP_H];
    for (int k = 0; k < 4; k++) xh[k] = (x[k] - norm[k]) * norm[4 + k];
    for (int j = 0; j < H; j++) {
        double z = b1[j];
        for (int k = 0; k < 4; k++) z += W1[j * 4 + k] * xh[k];
        h1[j] = tanh(z);
    }
    for (int j = 0; j < H; j++) {
        double z = b2[j];
        for (int k = 0; k < H; k++) z += W2[j * H + k] * h1[k];
        h2[j] = tanh(z);
    }
    double y = b3;
    for (int j = 0; j < H; j++) y += w3[j] * h2[j];
    if (g_out) *g_out = y;
    if (!grad) return 0.0;
    /* loss term (y a - t)^2; dy = 2 (y a - t) a (the 1/n is applied by the caller) */
    const double a = dy_scale_a, res = y * a - t, dy = 2.0 * res * a;
    double* gW1 = grad;
    double* gb1 = gW1 + H * 4;
    double* gW2 = gb1 + H;
    double* gb2 = gW2 + H * H;
    double* gw3 = gb2 + H;
    double d2[ORC_MLP_H], d1[ORC_MLP_H];
    gw3[H] += dy;
    for (int j = 0; j < H; j++) {
        gw3[j] += dy * h2[j];
        d2[j] = dy * w3[j] * (1.0 - h2[j] * h2[j]);
    }
    for (int j = 0; j < H; j++) {
        gb2[j] += d2[j];
        for (int k = 0; k < H; k++) gW2[j * H + k] += d2[j] * h1[k];
    }
    for (int k = 0; k < H; k++) {
        double s = 0.0;
        for (int j = 0; j < H; j++) s += W2[j * H + k] * d2[j];
        d1[k] = s * (1.0 - h1[k] * h1[k]);
    }
    for (int j = 0; j < H; j++) {
        gb1[j] += d1[j];
        for (int k = 0; k < 4; k++) gW1[j * 4 + k] += d1[j] * xh[k];
    }
    return res * res;
}

void orc_mlp_forward(const double* params, const double* norm, long n, const double* X, double* g)
{
    for (long i = 0; i < n; i++) mlp_one(params, norm, X + 4 * i, 0.0, 0.0, NULL, g + i);
}

double orc_mlp_loss_grad(const double* params, const double* norm, long n, const double* X, const double* a,
                         const double* t, double* grad)
{
    for (int p = 0; p < ORC_MLP_NP; p++) grad[p] = 0.0;
    double loss = 0.0;
    for (long i = 0; i < n; i++) loss += mlp_one(params, norm, X + 4 * i, a[i], t[i], grad, NULL);
    const double inv = n > 0 ? 1.0 / (double)n : 0.0;
    for (int p = 0; p < ORC_MLP_NP; p++) grad[p] *= inv;
    return loss * inv;
}
