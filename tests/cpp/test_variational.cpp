// C++ API test, mirroring TensorNetworksUnittests/VariationalGroundStateTests.C
// (the reference's gtest fixture builds H and the MPS through the Factory and
// checks E/(L-1) against recorded values).  Needs a HIP device.
#include <cmath>
#include <cstdio>
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
    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("all C++ API tests passed\n");
    return 0;
}
