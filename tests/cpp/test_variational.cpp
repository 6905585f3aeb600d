// C++ API test, mirroring TensorNetworksUnittests/VariationalGroundStateTests.C
// (the reference's gtest fixture builds H and the MPS through the Factory and
// checks E/(L-1) against recorded values).  Needs a HIP device.
#include <cmath>
#include <algorithm>
#include <cstdio>
#include <vector>
#include <memory>
#include <string>

#include "../../include/tn/TensorNetworks.H"

using namespace TensorNetworks;

static int failures = 0;
#define EXPECT_NEAR(a, b, tol)                                                                    \
    do {                                                                                          \
        double _a = (a), _b = (b);                                                                \
        if (!(std::fabs(_a - _b) <= (tol))) {                                                     \
            std::printf("FAIL %s:%d  %s = %.15g vs %.15g (tol %g)\n", __FILE__, __LINE__, #a, _a, \
                        _b, (double)(tol));                                                       \
            failures++;                                                                           \
        }                                                                                         \
    } while (0)
#define EXPECT_EQ(a, b)                                                                        \
    do {                                                                                       \
        if (!((a) == (b))) {                                                                   \
            std::printf("FAIL %s:%d  %s == %s\n", __FILE__, __LINE__, #a, #b);                 \
            failures++;                                                                        \
        }                                                                                      \
    } while (0)
#define EXPECT_LT(a, b)                                                                        \
    do {                                                                                       \
        if (!((a) < (b))) {                                                                    \
            std::printf("FAIL %s:%d  %s < %s\n", __FILE__, __LINE__, #a, #b);                  \
            failures++;                                                                        \
        }                                                                                      \
    } while (0)

struct Fixture {
    std::unique_ptr<Factory> f{Factory::GetFactory()};
    std::unique_ptr<Hamiltonian> H;
    std::unique_ptr<MPS> psi;
    void Setup(int L, double S, int D, MPOForm form)
    {
        H.reset(f->Make1D_NN_HeisenbergHamiltonian(L, S, form, 1.0, 1.0, 0.0));
        psi.reset(H->CreateMPS(D, 1e-12, 1e-12));
    }
};

// VariationalGroundStateTests.C:75-84
static void TestIdentityOperator()
{
    Fixture t;
    t.Setup(10, 0.5, 2, RegularLower);
    t.psi->InitializeWith(Random);
    t.psi->Normalize(DLeft);
    std::unique_ptr<MPO> IO(t.H->CreateUnitOperator());
    EXPECT_NEAR(t.psi->GetExpectation(IO.get()), 1.0, 1e-13);
}

// VariationalGroundStateTests.C:86-109
static void TestSweep_L2S1D2(MPOForm form)
{
    Fixture t;
    int L = 2, D = 2, maxIter = 100;
    t.Setup(L, 0.5, D, form);
    t.psi->InitializeWith(Random);
    Epsilons eps(1e-12);
    eps.itsDelatEnergy1Epsilon = 1e-9;
    IterationSchedule is;
    is.Insert({maxIter, D, eps});
    t.psi->FindVariationalGroundState(t.H.get(), is);
    double E1 = t.psi->GetExpectation(t.H.get());
    EXPECT_NEAR(E1 / (L - 1), -0.75, 1e-7);
    std::unique_ptr<MPO> H2(t.H->CreateH2Operator());
    double E2 = t.psi->GetExpectation(H2.get());
    EXPECT_EQ(H2->GetMaxDw(), 5);  // :106, Parker compression
    EXPECT_NEAR(E2, E1 * E1, 1e-13);
}

// VariationalGroundStateTests.C:137-153, :220-236
static void TestSweep_L9(int D, double Eexpect, double tol, MPOForm form)
{
    Fixture t;
    int L = 9, maxIter = 100;
    t.Setup(L, 0.5, D, form);
    t.psi->InitializeWith(Random);
    Epsilons eps(1e-12);
    eps.itsDelatEnergy1Epsilon = 1e-10;
    IterationSchedule is;
    is.Insert({maxIter, D, eps});
    t.psi->FindVariationalGroundState(t.H.get(), is);
    EXPECT_NEAR(t.psi->GetExpectation(t.H.get()) / (L - 1), Eexpect, tol);
    EXPECT_LT(t.psi->GetReport().nsweeps, maxIter + 1);
}

// PyTensorNetworks/test_Unittest1.py:22 (D grown 2 -> 8 in steps of 1)
static void TestGrowD_L10()
{
    Fixture t;
    int L = 10;
    t.Setup(L, 0.5, 2, RegularUpper);
    t.psi->InitializeWith(Random);
    Epsilons eps;  // 1e-12 everywhere
    IterationSchedule is;
    is.Insert({100, 8, eps});
    double de2 = t.psi->FindVariationalGroundState(t.H.get(), is);
    EXPECT_NEAR(t.psi->GetExpectation(t.H.get()) / (L - 1), -0.473113399629092, 1e-11);
    EXPECT_LT(de2, 5e-5);
}

// VariationalGroundStateTests.C:426-638: the other Factory models
static void TestModel(Hamiltonian* Hraw, int L, const std::vector<int>& Ds, double Eexpect, double tolE, int dwH2,
                      double tolE2, bool compress, int dwCanon, int dwComp)
{
    std::unique_ptr<Hamiltonian> H(Hraw);
    if (compress) {
        H->CanonicalForm();
        EXPECT_EQ(H->GetMaxDw(), dwCanon);
        H->Compress(Parker, 0, 1e-10);
        EXPECT_EQ(H->GetMaxDw(), dwComp);
    }
    std::unique_ptr<MPS> psi(H->CreateMPS(Ds[0], 1e-12, 1e-12));
    psi->InitializeWith(Random);
    Epsilons eps(1e-12);
    eps.itsDelatEnergy1Epsilon = 1e-9;
    eps.itsDeltaLambdaEpsilon = 1e-9;
    IterationSchedule is;
    for (int D : Ds) is.Insert({100, D, eps});
    psi->FindVariationalGroundState(H.get(), is);
    const double E1 = psi->GetExpectation(H.get());
    EXPECT_NEAR(E1 / (L - 1), Eexpect, tolE);
    std::unique_ptr<MPO> H2(H->CreateH2Operator());
    EXPECT_EQ(H2->GetMaxDw(), dwH2);
    EXPECT_NEAR(psi->GetExpectation(H2.get()), E1 * E1, tolE2);
}

// GpuHeffClient + GpuSparseEigenSolver (the SparseEigenSolver seam) against a
// dense eigensolve of the same operator, built column by column through
// DoMatVecContraction: both eigenpairs of LapackEigenSolver (IL=1, IU=2)
static void jacobi_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& Z)
{
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Z[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
        if (off < 1e-30) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A[(size_t)p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double th = (A[(size_t)q * n + q] - A[(size_t)p * n + p]) / (2 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double zkp = Z[(size_t)k * n + p], zkq = Z[(size_t)k * n + q];
                    Z[(size_t)k * n + p] = c * zkp - s * zkq;
                    Z[(size_t)k * n + q] = s * zkp + c * zkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; i++) w[i] = A[(size_t)i * n + i];
}

static void TestSparseEigenSolverSeam()
{
    tn_ctx* ctx = nullptr;
    if (tn_ctx_create(0, &ctx) != 0) { failures++; return; }
    const int d = 2, D1 = 3, D2 = 5, Dw = 3;
    std::vector<double> W((size_t)Dw * Dw * d * d), L((size_t)Dw * D1 * D1), R((size_t)Dw * D2 * D2);
    auto rnd = [](int k) { return std::sin(1.7 * k + 0.3 * k * k); };
    for (int w1 = 0; w1 < Dw; w1++)
        for (int w2 = 0; w2 < Dw; w2++)
            for (int m = 0; m < d; m++)
                for (int n = 0; n < d; n++) {
                    const int a = std::min(m, n), b = std::max(m, n);
                    W[((w1 * Dw + w2) * d + m) * d + n] = rnd(((w1 * Dw + w2) * d + a) * d + b);
                }
    for (int w = 0; w < Dw; w++)
        for (int i = 0; i < D1; i++)
            for (int j = 0; j < D1; j++) L[w + Dw * (i + D1 * j)] = rnd(100 + w * 50 + std::min(i, j) * 7 + std::max(i, j));
    for (int w = 0; w < Dw; w++)
        for (int i = 0; i < D2; i++)
            for (int j = 0; j < D2; j++) R[w + Dw * (i + D2 * j)] = rnd(900 + w * 50 + std::min(i, j) * 7 + std::max(i, j));
    SiteOperator so(d, Dw, Dw, W);
    GpuHeffClient client(ctx, so, D1, D2, L.data(), R.data());
    const int N = (int)client.GetSize();
    std::vector<double> H((size_t)N * N), e(N), y(N);
    for (int c = 0; c < N; c++) {
        std::fill(e.begin(), e.end(), 0.0);
        e[c] = 1.0;
        client.DoMatVecContraction(N, e.data(), y.data());
        for (int r = 0; r < N; r++) H[(size_t)r * N + c] = y[r];
    }
    double asym = 0;
    for (int r = 0; r < N; r++)
        for (int c = 0; c < N; c++) asym = std::fmax(asym, std::fabs(H[(size_t)r * N + c] - H[(size_t)c * N + r]));
    EXPECT_NEAR(asym, 0.0, 1e-13);  // LapackEigenSolver.C:71-77 Hermiticity
    std::vector<double> w, Z;
    jacobi_eig(N, H, w, Z);
    std::vector<int> ord(N);
    for (int i = 0; i < N; i++) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return w[a] < w[b]; });
    GpuSparseEigenSolver solver;
    auto [U, ev] = solver.Solve(&client, 1e-12, 2);
    EXPECT_EQ(U.GetNumCols(), 2);
    for (int k = 0; k < 2; k++) {
        EXPECT_NEAR(ev(k + 1), w[ord[k]], 1e-10);
        double ov = 0, nrm = 0;
        for (int i = 0; i < N; i++) {
            ov += U(i + 1, k + 1) * Z[(size_t)i * N + ord[k]];
            nrm += U(i + 1, k + 1) * U(i + 1, k + 1);
        }
        EXPECT_NEAR(std::fabs(ov), 1.0, 1e-8);
        EXPECT_NEAR(nrm, 1.0, 1e-12);
    }
    tn_ctx_destroy(ctx);
}

// GpuQRSolver behind the QRSolver seam (MPSSite::NormalizeQR's SolveThinQR / SolveThinLQ)
static void TestQRSeam()
{
    tn_ctx* ctx = nullptr;
    if (tn_ctx_create(0, &ctx) != 0) { failures++; return; }
    GpuQRSolver qr(ctx);
    MatrixRT A(7, 4), B(4, 7);
    for (int i = 1; i <= 7; i++)
        for (int j = 1; j <= 4; j++) {
            A(i, j) = std::sin(1.3 * i * j + 0.7 * i);
            B(j, i) = std::cos(0.9 * i * j + 0.2 * j);
        }
    auto [Q, R] = qr.SolveThinQR(A);
    double err = 0, orth = 0, low = 0;
    for (int i = 1; i <= 7; i++)
        for (int j = 1; j <= 4; j++) {
            double r = 0;
            for (int k = 1; k <= 4; k++) r += Q(i, k) * R(k, j);
            err = std::fmax(err, std::fabs(r - A(i, j)));
        }
    for (int a = 1; a <= 4; a++)
        for (int b = 1; b <= 4; b++) {
            double g = 0;
            for (int i = 1; i <= 7; i++) g += Q(i, a) * Q(i, b);
            orth = std::fmax(orth, std::fabs(g - (a == b)));
            if (a > b) low = std::fmax(low, std::fabs(R(a, b)));
        }
    EXPECT_NEAR(err, 0.0, 1e-13);
    EXPECT_NEAR(orth, 0.0, 1e-13);
    EXPECT_NEAR(low, 0.0, 0.0);
    auto [Lm, Q2] = qr.SolveThinLQ(B);
    err = 0;
    for (int i = 1; i <= 4; i++)
        for (int j = 1; j <= 7; j++) {
            double r = 0;
            for (int k = 1; k <= 4; k++) r += Lm(i, k) * Q2(k, j);
            err = std::fmax(err, std::fabs(r - B(i, j)));
        }
    EXPECT_NEAR(err, 0.0, 1e-13);
    tn_ctx_destroy(ctx);
}

// MPSNormTests.C:71-90 (NormalizeQR for the QR gauge) and checkpoint / resume
static void TestNormStatusAndCheckpoint()
{
    Fixture t;
    t.Setup(10, 0.5, 4, RegularUpper);
    t.psi->InitializeWith(Random);
    t.psi->Normalize(DLeft);
    EXPECT_EQ(t.psi->GetNormStatus(), std::string("AAAAAAAAAA"));
    t.psi->NormalizeQR(DRight);
    EXPECT_EQ(t.psi->GetNormStatus(), std::string("BBBBBBBBBB"));
    EXPECT_EQ(t.psi->GetNormStatus(3), 'B');
    const double E = t.psi->GetExpectation(t.H.get());
    const std::string path = "/tmp/tn_cpp_checkpoint.tnck";
    t.psi->Save(path, true);
    std::unique_ptr<MPS> back(MPS::Load(path));
    EXPECT_NEAR(back->GetExpectation(t.H.get()), E, 1e-13 * std::fabs(E));
    EXPECT_EQ(back->GetNormStatus(), std::string("BBBBBBBBBB"));
    std::remove(path.c_str());
}

// SVCompressorImp.C:19-73 through the Factory
static void TestCompressor()
{
    std::unique_ptr<Factory> f(Factory::GetFactory());
    std::unique_ptr<SVCompressorC> c(f->MakeMPSCompressor(2, 0.0));
    MatrixRT U(4, 4, 1.0), Vt(4, 3, 1.0);
    DiagonalMatrixRT s(4);
    s(1) = 0.8; s(2) = 0.5; s(3) = 0.3; s(4) = 0.1;
    double disc = c->Compress(U, s, Vt);
    EXPECT_NEAR(disc, 0.1, 1e-15);
    EXPECT_NEAR((double)s.size(), 2.0, 0);
    EXPECT_NEAR(s(1), 0.8 * (0.99 / 0.89), 1e-15);
    EXPECT_NEAR((double)U.GetNumCols(), 2.0, 0);
}

// GpuSVDSolver behind the SVDSolver seam
static void TestSVDSeam()
{
    tn_ctx* ctx = nullptr;
    if (tn_ctx_create(0, &ctx) != 0) { failures++; return; }
    GpuSVDSolver svd(ctx);
    MatrixRT A(5, 3);
    for (int i = 1; i <= 5; i++)
        for (int j = 1; j <= 3; j++) A(i, j) = std::sin(1.0 * i * j + 0.3 * i);
    auto [U, s, Vt] = svd.SolveAll(A, 1e-13);
    double err = 0;
    for (int i = 1; i <= 5; i++)
        for (int j = 1; j <= 3; j++) {
            double r = 0;
            for (int k = 1; k <= 3; k++) r += U(i, k) * s(k) * Vt(k, j);
            err = std::fmax(err, std::fabs(r - A(i, j)));
        }
    EXPECT_NEAR(err, 0.0, 1e-13);
    tn_ctx_destroy(ctx);
}

int main()
{
    TestIdentityOperator();
    TestSweep_L2S1D2(RegularLower);
    TestSweep_L2S1D2(RegularUpper);
    TestSweep_L9(2, -0.45317425, 1e-7, RegularLower);
    TestSweep_L9(2, -0.45317425, 1e-7, RegularUpper);
    TestSweep_L9(4, -0.46664265599414939, 1e-7, RegularLower);
    TestSweep_L9(4, -0.46664265599414939, 1e-7, RegularUpper);
    TestGrowD_L10();
    TestCompressor();
    TestSVDSeam();
    TestSparseEigenSolverSeam();
    TestQRSeam();
    TestNormStatusAndCheckpoint();
    {
        std::unique_ptr<Factory> f(Factory::GetFactory());
        // :426-449, :476-498, :500-523
        TestModel(f->Make1D_NN_TransverseIsingHamiltonian(2, 0.5, RegularLower, 1.0, 0.0), 2, {2}, -0.25, 1e-7, 2,
                  1e-14, false, 0, 0);
        TestModel(f->Make1D_NN_TransverseIsingHamiltonian(9, 0.5, RegularLower, 1.0, 1.0), 9, {2}, -0.594125875,
                  1e-7, 5, 1e-4, false, 0, 0);
        TestModel(f->Make1D_NN_TransverseIsingHamiltonian(9, 0.5, RegularUpper, 1.0, 1.0), 9, {2, 8}, -0.5941321968,
                  1e-7, 5, 1e-12, false, 0, 0);
        // :550-580 (three-body, compressed to Dw = 3), :581-609 (two-body long range, NN = 3)
        TestModel(f->Make1D_3BodyHamiltonian(9, 0.5, RegularUpper, 1.0, 1.0, 0.5), 9, {2, 4}, -0.3471737999, 1e-7, 5,
                  1e-5, true, 5, 3);
        TestModel(f->Make1D_2BodyLongRangeHamiltonian(9, 0.5, RegularLower, 1.0, 0.0, 3), 9, {2, 2}, -0.21614583333,
                  1e-7, 12, 3e-14, true, 8, 5);
        // three-body long range: exact operator, no recorded energy; check the
        // variance identity on a converged state instead
        std::unique_ptr<Hamiltonian> H3(f->Make1D_3BodyLongRangeHamiltonian(6, 0.5, RegularUpper, 1.0, 0.0, 2));
        EXPECT_EQ(H3->GetMaxDw(), 5);
    }
    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("all C++ API tests passed\n");
    return 0;
}
