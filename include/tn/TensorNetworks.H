// TensorNetworks.H -- C++ API of the MI355X DMRG path, mirroring the public
// interface of JanReimers/TensorNetworks (TensorNetworks/*.H) for the
// variational ground-state sweep.  Every class here is host code over the C ABI
// in include/tn_abi.h; the arithmetic runs in the gfx950 kernels.
//
//   Factory                TensorNetworks/Factory.H:9-33
//   Hamiltonian            TensorNetworks/Hamiltonian.H:15-41
//   MPO / SiteOperator     TensorNetworks/MPO.H:17-44, SiteOperator.H:20-81
//   MPS                    TensorNetworks/MPS.H:14-75
//   IterationSchedule      TensorNetworks/IterationSchedule.H:13-59
//   Epsilons               TensorNetworks/Epsilons.H:11-26
//   Enums                  TensorNetworks/Enums.H:8-16
//   SVCompressor           TensorNetworks/SVCompressor.H:10-24
//   TNSLogger              TensorNetworks/TNSLogger.H:14-50
//   EigenSolver / SparseEigenSolver / SVDSolver / QRSolver  NumericalMethods/*.H (solver seam)
#ifndef TN_TENSORNETWORKS_H
#define TN_TENSORNETWORKS_H

#include <algorithm>
#include <cstdarg>
#include <string>
#include <tuple>
#include <vector>

#include "../tn_abi.h"

namespace TensorNetworks {

// ---- Enums.H ----------------------------------------------------------------
enum Position { PLeft, PBulk, PRight };
enum Direction { DLeft, DRight };
enum State { Product, Random, Neel, Constant };
enum MPOForm { FUnknown, RegularUpper, RegularLower, expH, FUnit };
enum CompressType { CNone, Std, Parker };
// arithmetic of the contractions (extension: the reference is FP64 complex only)
enum Precision { FP64 = TN_FP64, FP32 = TN_FP32, BF16 = TN_BF16 };

// ---- dense containers (column-major, 1-based like the reference's oml) -----
template <class T> class Vector {
public:
    Vector() {}
    explicit Vector(int n, T v = T()) : d_(n, v) {}
    int size() const { return (int)d_.size(); }
    T& operator()(int i) { return d_[i - 1]; }
    const T& operator()(int i) const { return d_[i - 1]; }
    T* data() { return d_.data(); }
    const T* data() const { return d_.data(); }
private:
    std::vector<T> d_;
};

template <class T> class Matrix {
public:
    Matrix() : m_(0), n_(0) {}
    Matrix(int m, int n, T v = T()) : m_(m), n_(n), d_((size_t)m * n, v) {}
    int GetNumRows() const { return m_; }
    int GetNumCols() const { return n_; }
    T& operator()(int i, int j) { return d_[(size_t)(j - 1) * m_ + (i - 1)]; }
    const T& operator()(int i, int j) const { return d_[(size_t)(j - 1) * m_ + (i - 1)]; }
    T* data() { return d_.data(); }
    const T* data() const { return d_.data(); }
private:
    int m_, n_;
    std::vector<T> d_;
};
typedef Matrix<double> MatrixRT;
typedef Vector<double> VectorRT;
typedef Vector<double> DiagonalMatrixRT;  // singular values (diagonal stored as a vector)

// ---- Epsilons.H / Epsilons.C ------------------------------------------------
struct Epsilons {
    explicit Epsilons(double default_eps = 1e-12);
    double itsDelatEnergy1Epsilon;
    double itsDelatEnergy2Epsilon;
    double itsDelatNormEpsilon;
    double itsEigenSolverEpsilon;
    double itsNormalizationEpsilon;
    double itsMPSCompressEpsilon;
    double itsMPOCompressEpsilon;
    double itsSparseMatrixEpsilon;
    double itsDeltaLambdaEpsilon;
};

// ---- IterationSchedule.H ----------------------------------------------------
struct IterationScheduleLine {
    IterationScheduleLine(int maxiter, int Dmax, const Epsilons&);
    IterationScheduleLine(int maxiter, int Dmax, int DeltaD, const Epsilons&);
    int itsMaxGSSweepIterations;
    int itsMaxOptimizeIterations;
    int itsDmax;
    int itsDeltaD;
    double itsdt;
    Epsilons itsEps;
};

class IterationSchedule {
public:
    IterationSchedule() : itsCurrentLine(0) {}
    void Insert(const IterationScheduleLine& l) { itsLines.push_back(l); }
    void begin() const { itsCurrentLine = 0; }
    bool end() const { return itsCurrentLine == itsLines.size(); }
    // moves to the next line and returns it (IterationSchedule.H:45-49); past
    // the last line it returns the last line instead of reading beyond the end
    const IterationScheduleLine& operator++(int) const
    {
        itsCurrentLine++;
        return itsLines[std::min<size_t>(itsCurrentLine, itsLines.size() - 1)];
    }
    const IterationScheduleLine& operator*() const { return itsLines[itsCurrentLine]; }
private:
    mutable unsigned int itsCurrentLine;
    std::vector<IterationScheduleLine> itsLines;
};

// ---- TNSLogger.H ------------------------------------------------------------
class TNSLogger {
public:
    explicit TNSLogger(int level = 0) : itsLevel(level) {}
    virtual ~TNSLogger() {}
    virtual void LogInfo(int level, const std::string& msg);
    virtual void LogInfoV(int level, const char* fmt, ...);
protected:
    int itsLevel;
};
extern TNSLogger* Logger;  // global logger (NullLogger-like default: level -1)

// ---- SiteOperator / MPO -----------------------------------------------------
// Operator-valued matrix of one site: Dw1 x Dw2 blocks of d x d reals,
// W[((w1*Dw2 + w2)*d + m)*d + n], m bra, n ket.
class SiteOperator {
public:
    SiteOperator(int d, int Dw1, int Dw2, std::vector<double> W)
        : itsd(d), itsDw1(Dw1), itsDw2(Dw2), itsW(std::move(W)) {}
    int Getd() const { return itsd; }
    int GetDw1() const { return itsDw1; }
    int GetDw2() const { return itsDw2; }
    const std::vector<double>& GetW() const { return itsW; }
    double W(int w1, int w2, int m, int n) const { return itsW[((w1 * itsDw2 + w2) * itsd + m) * itsd + n]; }
private:
    int itsd, itsDw1, itsDw2;
    std::vector<double> itsW;
};

class MPO {
public:
    MPO() {}
    explicit MPO(std::vector<SiteOperator> ops, MPOForm f = FUnknown) : itsOps(std::move(ops)), itsForm(f) {}
    virtual ~MPO() {}
    int GetL() const { return (int)itsOps.size(); }
    const SiteOperator* GetSiteOperator(int isite) const { return &itsOps[isite - 1]; }  // 1-based
    int GetMaxDw() const;
    MPOForm GetForm() const { return itsForm; }
    // MPO::Product (TensorNetworks/MPO.C:73-79), uncompressed: a new MPO
    MPO* Product(const MPO* other) const;
    // MPO::CanonicalForm (MPO.C:120-128): block QR sweeps left then right;
    // needs a regular form (RegularUpper / RegularLower)
    void CanonicalForm();
    // MPO::Compress(CompressType, Dmax, eps) (MPO.C:93-118): Parker compression
    // of the interior channels (Std is not provided: throws); returns the
    // truncation error per bond
    double Compress(CompressType, int Dmax, double epsSV);
protected:
    std::vector<SiteOperator> itsOps;
    MPOForm itsForm = FUnknown;
};

class MPS;

// ---- Hamiltonian.H ----------------------------------------------------------
class Hamiltonian : public MPO {
public:
    // nearest-neighbour XXZ chain (Hamiltonian_1D_NN_Heisenberg.C:35-79) + Dz Sz^2
    Hamiltonian(int L, double S, MPOForm f, double Jxy, double Jz, double hz, double Dz);
    // any model from its bulk operator-valued matrix Wb (Dw x Dw blocks of d x d,
    // [((w1*Dw + w2)*d + m)*d + n]), edges kept as SiteOperatorImp::Init_lr does
    Hamiltonian(int L, double S, MPOForm f, int Dw, const std::vector<double>& Wb);
    int Getd() const { return itsd; }
    double GetS() const { return itsS; }
    MPS* CreateMPS(int D, double normEps = 1e-12, double epsSV = 1e-12) const;
    MPO* CreateUnitOperator() const;
    // Hamiltonian.C:8-16: H*H, CanonicalForm, Compress(Parker, 0, 1e-13)
    MPO* CreateH2Operator() const;
private:
    int itsL, itsd;
    double itsS;
};

// ---- SVCompressor.H ---------------------------------------------------------
template <class T> class SVCompressor {
public:
    SVCompressor(int Dmax, double eps) : itsDmax(Dmax), itsSVeps(eps) {}
    virtual ~SVCompressor() {}
    int GetDmax() const { return itsDmax; }
    bool Donly() const { return itsSVeps == 0.0; }
    // SVCompressorImp.C:19-73: returns the discarded weight sum(s^2)
    double Compress(Matrix<T>& U, DiagonalMatrixRT& s, Matrix<T>& Vdagger) const;
private:
    int itsDmax;
    double itsSVeps;
};
typedef SVCompressor<double> SVCompressorR;
typedef SVCompressor<double> SVCompressorC;  // real arithmetic on this path

// ---- solver seam (NumericalMethods/*.H) -------------------------------------
template <class T> class SparseEigenSolverClient {
public:
    virtual ~SparseEigenSolverClient() {}
    virtual long int GetSize() const = 0;
    virtual void DoMatVecContraction(int N, const T* xvec, T* yvec) const = 0;
};

template <class T> class SparseEigenSolver {
public:
    typedef std::tuple<Matrix<T>, Vector<double>> UdType;
    virtual ~SparseEigenSolver() {}
    virtual void Reset() = 0;
    virtual UdType Solve(const SparseEigenSolverClient<T>*, double eps, int NumEigenValues) = 0;
};

// QR seam (NumericalMethods/QRSolver.H; LapackQRSolver.C): thin QR (A = Q R,
// m >= n) and thin LQ (A = L Q, m <= n), as MPSSite::NormalizeQR uses them
template <class T> class QRSolver {
public:
    typedef std::tuple<Matrix<T>, Matrix<T>> QRType;
    virtual ~QRSolver() {}
    virtual QRType SolveThinQR(const Matrix<T>& A) = 0;  // returns (Q, R)
    virtual QRType SolveThinLQ(const Matrix<T>& A) = 0;  // returns (L, Q)
};

template <class T> class SVDSolver {
public:
    typedef std::tuple<Matrix<T>, DiagonalMatrixRT, Matrix<T>> UsVType;
    virtual ~SVDSolver() {}
    virtual UsVType Solve(const Matrix<T>&, double eps, int NumSingularValues) = 0;
    virtual UsVType SolveAll(const Matrix<T>& A, double eps)
    {
        return Solve(A, eps, std::min(A.GetNumRows(), A.GetNumCols()));
    }
};

// One site's effective Hamiltonian on the GPU: the matrix-free client that
// replaces MPSSite::GetHeff (MPSSite_Contractions.C:155-193).
class GpuHeffClient : public SparseEigenSolverClient<double> {
public:
    GpuHeffClient(tn_ctx* ctx, const SiteOperator& W, int D1, int D2, const double* Lenv,
                  const double* Renv, Precision p = FP64);
    long int GetSize() const override { return itsN; }
    void DoMatVecContraction(int N, const double* x, double* y) const override;
    tn_ctx* ctx() const { return itsCtx; }
private:
    tn_ctx* itsCtx;
    long itsN;
};

// Device Lanczos behind the SparseEigenSolver seam (replaces LapackEigenSolver /
// PrimeEigenSolver for the site problem).  Accepts GpuHeffClient only: there is
// no host fallback.
class GpuSparseEigenSolver : public SparseEigenSolver<double> {
public:
    void Reset() override { itsGuess.clear(); }
    UdType Solve(const SparseEigenSolverClient<double>*, double eps, int NumEigenValues) override;
private:
    std::vector<double> itsGuess;
};

// Device CholeskyQR behind the QRSolver seam (replaces LapackQRSolver).
class GpuQRSolver : public QRSolver<double> {
public:
    explicit GpuQRSolver(tn_ctx* ctx) : itsCtx(ctx) {}
    QRType SolveThinQR(const Matrix<double>& A) override;
    QRType SolveThinLQ(const Matrix<double>& A) override;
private:
    tn_ctx* itsCtx;
};

// Device one-sided Jacobi SVD behind the SVDSolver seam (replaces LapackSVDSolver).
class GpuSVDSolver : public SVDSolver<double> {
public:
    explicit GpuSVDSolver(tn_ctx* ctx) : itsCtx(ctx) {}
    UsVType Solve(const Matrix<double>&, double eps, int NumSingularValues) override;
private:
    tn_ctx* itsCtx;
};

// ---- MPS.H --------------------------------------------------------------------
struct SweepReport {
    tn_sweep_stats stats{};
    int nsweeps = 0;
};

class MPS {
public:
    MPS(const Hamiltonian* H, int D, double normEps, double epsSV, int Dcap = 0);
    ~MPS();
    MPS(const MPS&) = delete;  // owns device memory
    MPS& operator=(const MPS&) = delete;
    void InitializeWith(State, uint64_t seed = 20261202);
    void Normalize(Direction LR);
    // MPS::NormalizeQR (MPSSite_Normalization.C:80-107): QR gauge moves
    void NormalizeQR(Direction LR);
    // MPS::GetNormStatus (MPSImp.C:180-192, MPSSite.C:173-211) with normEps
    std::string GetNormStatus() const;
    char GetNormStatus(int isite) const;  // 1-based
    // checkpoint / resume (tn_chain_save / tn_chain_load): the state, its
    // bond values and optionally the environment caches
    void Save(const std::string& path, bool withEnvironments = false) const;
    static MPS* Load(const std::string& path, int Dcap = 0);
    // MPSImp::FindVariationalGroundState (MPSImp_Variational.C:16-62); returns <H^2>-<H>^2
    double FindVariationalGroundState(const Hamiltonian* H, const IterationSchedule&);
    double GetExpectation(const MPO* o) const;
    int GetMaxD() const;
    int GetL() const { return itsL; }
    // options of the local update (precision, Lanczos tolerance ...): extension
    void SetSweepOptions(const tn_sweep_opts& o) { itsOpts = o; }
    const SweepReport& GetReport() const { return itsReport; }
    tn_chain* chain() const { return itsChain; }
    tn_ctx* ctx() const { return itsCtx; }
private:
    double FindVariationalGroundState(const Hamiltonian* H, const IterationScheduleLine&);
    void Sweep(Direction lr, const Epsilons& eps);
    MPS() : itsL(0), itsd(0), itsNormEps(1e-12), itsCtx(nullptr), itsChain(nullptr) {}
    int itsL, itsd;
    double itsNormEps;
    tn_ctx* itsCtx;
    tn_chain* itsChain;
    tn_sweep_opts itsOpts;
    SweepReport itsReport;
};

// ---- Factory.H ----------------------------------------------------------------
class Factory {
public:
    static Factory* GetFactory();
    virtual ~Factory() {}
    Hamiltonian* Make1D_NN_HeisenbergHamiltonian(int L, double S, MPOForm, double Jxy, double Jz, double hz);
    // extension: single-ion anisotropy Dz*Sz^2 (the S=1 workload)
    Hamiltonian* Make1D_NN_HeisenbergHamiltonian(int L, double S, MPOForm, double Jxy, double Jz, double hz,
                                                 double Dz);
    // Factory.H:17-20 (Hamiltonian_1D_NN_TransverseIsing.C, Hamiltonian_1D_2Body_LongRange.C,
    // Hamiltonian_1D_3Body_LongRange.C, Hamiltonian_1D_3Body.C)
    Hamiltonian* Make1D_NN_TransverseIsingHamiltonian(int L, double S, MPOForm, double J, double hx);
    Hamiltonian* Make1D_2BodyLongRangeHamiltonian(int L, double S, MPOForm, double J, double hx, int NN);
    Hamiltonian* Make1D_3BodyLongRangeHamiltonian(int L, double S, MPOForm, double J, double hx, int NN);
    Hamiltonian* Make1D_3BodyHamiltonian(int L, double S, MPOForm, double J, double K, double hz);
    SVCompressorC* MakeMPSCompressor(int Dmax, double epsSV);
    TNSLogger* MakeSPDLogger(int level);
};

// status check helper: throws std::runtime_error with tn_last_error on failure
void check(tn_ctx* ctx, int rc, const char* what);

}  // namespace TensorNetworks

#endif
