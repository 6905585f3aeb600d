// tensornetworks.cpp -- C++ API (include/tn/TensorNetworks.H) over the C ABI.
#include "../../../include/tn/TensorNetworks.H"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace TensorNetworks {

void check(tn_ctx* ctx, int rc, const char* what)
{
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + (ctx ? tn_last_error(ctx) : "error"));
}

// ---- Epsilons.C:9-19 ---------------------------------------------------------
Epsilons::Epsilons(double e)
    : itsDelatEnergy1Epsilon(e), itsDelatEnergy2Epsilon(e), itsDelatNormEpsilon(e),
      itsEigenSolverEpsilon(e), itsNormalizationEpsilon(e), itsMPSCompressEpsilon(e),
      itsMPOCompressEpsilon(e), itsSparseMatrixEpsilon(e), itsDeltaLambdaEpsilon(e)
{
}

// ---- IterationSchedule.C:8-37 ------------------------------------------------
IterationScheduleLine::IterationScheduleLine(int maxiter, int Dmax, const Epsilons& eps)
    : IterationScheduleLine(maxiter, Dmax, 1, eps)
{
}
IterationScheduleLine::IterationScheduleLine(int maxiter, int Dmax, int DeltaD, const Epsilons& eps)
    : itsMaxGSSweepIterations(maxiter), itsMaxOptimizeIterations(0), itsDmax(Dmax), itsDeltaD(DeltaD),
      itsdt(0.0), itsEps(eps)
{
    if (maxiter <= 0 || Dmax <= 0 || DeltaD <= 0) throw std::invalid_argument("IterationScheduleLine");
}

// ---- logger --------------------------------------------------------------------
static TNSLogger theNullLogger(-1);
TNSLogger* Logger = &theNullLogger;

void TNSLogger::LogInfo(int level, const std::string& msg)
{
    if (level <= itsLevel) std::fprintf(stderr, "%s\n", msg.c_str());
}
void TNSLogger::LogInfoV(int level, const char* fmt, ...)
{
    if (level > itsLevel) return;
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(stderr, fmt, ap);
    va_end(ap);
    std::fprintf(stderr, "\n");
}

// ---- MPO --------------------------------------------------------------------------
int MPO::GetMaxDw() const
{
    int m = 0;
    for (const auto& o : itsOps) m = std::max(m, std::max(o.GetDw1(), o.GetDw2()));
    return m;
}

static void mpo_arrays(const MPO* o, std::vector<int>& dw1, std::vector<int>& dw2, std::vector<double>& w);

void MPO::CanonicalForm()
{
    if (itsForm != RegularUpper && itsForm != RegularLower)
        throw std::invalid_argument("MPO::CanonicalForm needs a regular (upper or lower) MPO");
    Compress(CNone, 0, 0.0);
}

double MPO::Compress(CompressType ct, int Dmax, double eps)
{
    // ct == CNone is used by CanonicalForm (block QR sweeps only)
    if (ct == Std) throw std::invalid_argument("MPO::Compress: only Parker compression is provided");
    if (itsForm != RegularUpper && itsForm != RegularLower)
        throw std::invalid_argument("MPO::Compress needs a regular (upper or lower) MPO");
    const int L = GetL(), d = itsOps[0].Getd();
    std::vector<int> dw1, dw2, o1(L), o2(L);
    std::vector<double> w;
    mpo_arrays(this, dw1, dw2, w);
    const int form = itsForm == RegularUpper ? 0 : 1, ops = ct == CNone ? 1 : 2;
    double te = 0.0;
    const int need = tn_mpo_reduce(L, d, dw1.data(), dw2.data(), w.data(), form, ops, Dmax, eps, o1.data(), o2.data(),
                                   nullptr, 0, &te);
    if (need < 0) throw std::invalid_argument("MPO::Compress: MPO not in regular form");
    std::vector<double> out(need);
    if (tn_mpo_reduce(L, d, dw1.data(), dw2.data(), w.data(), form, ops, Dmax, eps, o1.data(), o2.data(), out.data(),
                      need, &te) != need)
        throw std::runtime_error("MPO::Compress: inconsistent sizes");
    std::vector<SiteOperator> ops_new;
    size_t off = 0;
    for (int i = 0; i < L; i++) {
        const size_t n = (size_t)o1[i] * o2[i] * d * d;
        ops_new.emplace_back(d, o1[i], o2[i], std::vector<double>(out.begin() + off, out.begin() + off + n));
        off += n;
    }
    itsOps = std::move(ops_new);
    return te;
}

MPO* MPO::Product(const MPO* other) const
{
    // W2[(w1,v1),(w2,v2)](m,n) = sum_k W(w1,w2)(m,k) V(v1,v2)(k,n)
    std::vector<SiteOperator> ops;
    for (int i = 0; i < GetL(); i++) {
        const SiteOperator& a = itsOps[i];
        const SiteOperator& b = other->itsOps[i];
        const int d = a.Getd(), A1 = a.GetDw1(), A2 = a.GetDw2(), B1 = b.GetDw1(), B2 = b.GetDw2();
        std::vector<double> W((size_t)A1 * B1 * A2 * B2 * d * d, 0.0);
        for (int w1 = 0; w1 < A1; w1++)
            for (int v1 = 0; v1 < B1; v1++)
                for (int w2 = 0; w2 < A2; w2++)
                    for (int v2 = 0; v2 < B2; v2++)
                        for (int m = 0; m < d; m++)
                            for (int n = 0; n < d; n++) {
                                double acc = 0;
                                for (int k = 0; k < d; k++) acc += a.W(w1, w2, m, k) * b.W(v1, v2, k, n);
                                W[((((size_t)(w1 * B1 + v1) * (A2 * B2) + (w2 * B2 + v2)) * d + m) * d) + n] = acc;
                            }
        ops.emplace_back(d, A1 * B1, A2 * B2, std::move(W));
    }
    // the product of two regular MPOs of one form is regular in that form
    // (channel (w, v) -> w*Dw + v keeps the first and the last channel)
    return new MPO(std::move(ops), itsForm == other->itsForm ? itsForm : FUnknown);
}

// ---- Hamiltonian (Hamiltonian_1D_NN_Heisenberg.C:35-79, SiteOperatorImp.C:33-134) ----
// ---- spin operators (SpinCalculator.C:20-70): index n <-> Sz = n - S --------------
namespace {
struct SpinOps {
    int d;
    std::vector<double> Sp, Sm, Sz, Sx, I;
    explicit SpinOps(double S) : d((int)std::lround(2 * S + 1))
    {
        Sp.assign(d * d, 0.0);
        Sm.assign(d * d, 0.0);
        Sz.assign(d * d, 0.0);
        Sx.assign(d * d, 0.0);
        I.assign(d * d, 0.0);
        for (int m = 0; m < d; m++)
            for (int n = 0; n < d; n++) {
                const double sm = m - S, sn = n - S;
                if (m == n + 1) Sp[m * d + n] = std::sqrt(S * (S + 1) - sm * sn);
                if (m + 1 == n) Sm[m * d + n] = std::sqrt(S * (S + 1) - sm * sn);
                if (m == n) {
                    Sz[m * d + n] = sm;
                    I[m * d + n] = 1.0;
                }
            }
        for (int k = 0; k < d * d; k++) Sx[k] = 0.5 * (Sp[k] + Sm[k]);
    }
    std::vector<double> mul(const std::vector<double>& a, const std::vector<double>& b) const
    {
        std::vector<double> c(d * d, 0.0);
        for (int i = 0; i < d; i++)
            for (int k = 0; k < d; k++)
                for (int j = 0; j < d; j++) c[i * d + j] += a[i * d + k] * b[k * d + j];
        return c;
    }
};
// bulk operator-valued matrix under construction: block (a, b) = c * M
struct BulkW {
    int Dw, d;
    std::vector<double> W;
    BulkW(int Dw_, int d_) : Dw(Dw_), d(d_), W((size_t)Dw_ * Dw_ * d_ * d_, 0.0) {}
    void set(int a, int b, const std::vector<double>& M, double c = 1.0)
    {
        for (int k = 0; k < d * d; k++) W[(size_t)(a * Dw + b) * d * d + k] = c * M[k];
    }
};
}  // namespace

Hamiltonian::Hamiltonian(int L, double S, MPOForm f, double Jxy, double Jz, double hz, double Dz)
    : Hamiltonian(L, S, f, 5, [&]() {
          SpinOps o(S);
          const int d = o.d;
          std::vector<double> On(d * d, 0.0);
          for (int m = 0; m < d; m++) On[m * d + m] = hz * (m - S) + Dz * (m - S) * (m - S);
          BulkW w(5, d);
          if (f == RegularUpper) {
              w.set(0, 0, o.I); w.set(0, 1, o.Sp); w.set(0, 2, o.Sm); w.set(0, 3, o.Sz); w.set(0, 4, On);
              w.set(1, 4, o.Sm, 0.5 * Jxy); w.set(2, 4, o.Sp, 0.5 * Jxy); w.set(3, 4, o.Sz, Jz); w.set(4, 4, o.I);
          } else {
              w.set(0, 0, o.I); w.set(1, 0, o.Sp); w.set(2, 0, o.Sm); w.set(3, 0, o.Sz); w.set(4, 0, On);
              w.set(4, 1, o.Sm, 0.5 * Jxy); w.set(4, 2, o.Sp, 0.5 * Jxy); w.set(4, 3, o.Sz, Jz); w.set(4, 4, o.I);
          }
          return w.W;
      }())
{
}

Hamiltonian::Hamiltonian(int L, double S, MPOForm f, int Dw, const std::vector<double>& Wb)
    : itsL(L), itsd((int)std::lround(2 * S + 1)), itsS(S)
{
    if (L < 1 || itsd < 2) throw std::invalid_argument("Hamiltonian: L >= 1 and S >= 1/2");
    if (f != RegularUpper && f != RegularLower) throw std::invalid_argument("Hamiltonian: MPO form");
    const int d = itsd;
    if (Dw < 1 || Wb.size() != (size_t)Dw * Dw * d * d) throw std::invalid_argument("Hamiltonian: bulk W size");
    itsForm = f;
    const int lindex = (f == RegularUpper) ? 0 : Dw - 1, rindex = (f == RegularUpper) ? Dw - 1 : 0;
    for (int i = 0; i < L; i++) {
        int r0 = 0, r1 = Dw, c0 = 0, c1 = Dw;
        if (i == 0) { r0 = lindex; r1 = lindex + 1; }
        if (i == L - 1) { c0 = rindex; c1 = rindex + 1; }
        const int D1 = r1 - r0, D2 = c1 - c0;
        std::vector<double> W((size_t)D1 * D2 * d * d);
        for (int a = r0; a < r1; a++)
            for (int b = c0; b < c1; b++)
                for (int k = 0; k < d * d; k++)
                    W[(size_t)((a - r0) * D2 + (b - c0)) * d * d + k] = Wb[(size_t)(a * Dw + b) * d * d + k];
        itsOps.emplace_back(d, D1, D2, std::move(W));
    }
}

MPS* Hamiltonian::CreateMPS(int D, double normEps, double epsSV) const { return new MPS(this, D, normEps, epsSV); }

MPO* Hamiltonian::CreateUnitOperator() const
{
    std::vector<SiteOperator> ops;
    for (int i = 0; i < itsL; i++) {
        std::vector<double> W(itsd * itsd, 0.0);
        for (int m = 0; m < itsd; m++) W[m * itsd + m] = 1.0;
        ops.emplace_back(itsd, 1, 1, std::move(W));
    }
    return new MPO(std::move(ops));
}

MPO* Hamiltonian::CreateH2Operator() const
{
    MPO* H2 = Product(this);
    H2->CanonicalForm();
    H2->Compress(Parker, 0, 1e-13);
    return H2;
}

// ---- SVCompressorImp.C:19-73 ---------------------------------------------------------
template <class T> double SVCompressor<T>::Compress(Matrix<T>& U, DiagonalMatrixRT& s, Matrix<T>& Vd) const
{
    const int N = s.size();
    double disc = 0, resc = 1;
    const int D = tn_svcompress_rank(N, s.data(), itsDmax, itsSVeps, &disc, &resc);
    if (D < 0) throw std::runtime_error("SVCompressor: bad singular values");
    if (D < N) {
        Matrix<T> U2(U.GetNumRows(), D), V2(D, Vd.GetNumCols());
        DiagonalMatrixRT s2(D);
        for (int j = 1; j <= D; j++) {
            s2(j) = s(j) * resc;
            for (int i = 1; i <= U.GetNumRows(); i++) U2(i, j) = U(i, j);
            for (int k = 1; k <= Vd.GetNumCols(); k++) V2(j, k) = Vd(j, k);
        }
        U = U2;
        Vd = V2;
        s = s2;
    }
    return disc;
}
template class SVCompressor<double>;

// ---- solver seam -------------------------------------------------------------------------
GpuHeffClient::GpuHeffClient(tn_ctx* ctx, const SiteOperator& W, int D1, int D2, const double* L,
                             const double* R, Precision p)
    : itsCtx(ctx), itsN((long)W.Getd() * D1 * D2)
{
    check(ctx, tn_site_bind(ctx, W.Getd(), D1, D2, W.GetDw1(), W.GetDw2(), W.GetW().data(), L, R, (int)p),
          "tn_site_bind");
}

void GpuHeffClient::DoMatVecContraction(int N, const double* x, double* y) const
{
    check(itsCtx, tn_heff_apply(itsCtx, x, y, N), "tn_heff_apply");
}

SparseEigenSolver<double>::UdType GpuSparseEigenSolver::Solve(const SparseEigenSolverClient<double>* cl,
                                                             double eps, int nev)
{
    const GpuHeffClient* hc = dynamic_cast<const GpuHeffClient*>(cl);
    if (!hc) throw std::invalid_argument("GpuSparseEigenSolver: client must be a GpuHeffClient (no host fallback)");
    if (nev < 1 || nev > 2) throw std::invalid_argument("GpuSparseEigenSolver: 1 <= NumEigenValues <= 2");
    const long N = hc->GetSize();
    if (nev == 2 && N < 2) throw std::invalid_argument("GpuSparseEigenSolver: two eigenpairs need N >= 2");
    if ((long)itsGuess.size() != N) itsGuess.assign(N, 0.0);
    double ev[2] = {0, 0};
    int nmv = 0;
    std::vector<double> X((size_t)nev * N);
    std::copy(itsGuess.begin(), itsGuess.end(), X.begin());  // warm start (Reset() clears it)
    const int rc = tn_eig_lowest(hc->ctx(), N, nev, eps, 200, ev, X.data(), &nmv);
    check(hc->ctx(), rc, "tn_eig_lowest");
    std::copy(X.begin(), X.begin() + N, itsGuess.begin());
    Matrix<double> U(N, nev, 0.0);
    for (int k = 0; k < nev; k++)
        for (long i = 0; i < N; i++) U(i + 1, k + 1) = X[(size_t)k * N + i];  // column k+1: eigenvector k+1
    Vector<double> d(nev);
    for (int k = 0; k < nev; k++) d(k + 1) = ev[k];
    return std::make_tuple(U, d);
}

QRSolver<double>::QRType GpuQRSolver::SolveThinQR(const Matrix<double>& A)
{
    const int m = A.GetNumRows(), n = A.GetNumCols();
    Matrix<double> Q(m, n), R(n, n);
    const int rc = tn_qr(itsCtx, 0, m, n, A.data(), Q.data(), R.data());
    check(itsCtx, rc, "tn_qr");
    if (rc != 0) std::fprintf(stderr, "GpuQRSolver: orthogonality of Q above 1e-11\n");
    return std::make_tuple(Q, R);
}

QRSolver<double>::QRType GpuQRSolver::SolveThinLQ(const Matrix<double>& A)
{
    const int m = A.GetNumRows(), n = A.GetNumCols();
    Matrix<double> Lm(m, m), Q(m, n);
    const int rc = tn_qr(itsCtx, 1, m, n, A.data(), Q.data(), Lm.data());
    check(itsCtx, rc, "tn_qr");
    if (rc != 0) std::fprintf(stderr, "GpuQRSolver: orthogonality of Q above 1e-11\n");
    return std::make_tuple(Lm, Q);
}

SVDSolver<double>::UsVType GpuSVDSolver::Solve(const Matrix<double>& A, double, int ns)
{
    const int m = A.GetNumRows(), n = A.GetNumCols(), k = std::min(m, n);
    if (ns < 1 || ns > k) throw std::invalid_argument("GpuSVDSolver: NumSingularValues");
    Matrix<double> U(m, k), Vt(k, n);
    DiagonalMatrixRT s(k);
    check(itsCtx, tn_svd(itsCtx, m, n, A.data(), U.data(), s.data(), Vt.data()), "tn_svd");
    if (ns < k) {
        Matrix<double> U2(m, ns), V2(ns, n);
        DiagonalMatrixRT s2(ns);
        for (int j = 1; j <= ns; j++) {
            s2(j) = s(j);
            for (int i = 1; i <= m; i++) U2(i, j) = U(i, j);
            for (int c = 1; c <= n; c++) V2(j, c) = Vt(j, c);
        }
        return std::make_tuple(U2, s2, V2);
    }
    return std::make_tuple(U, s, Vt);
}

// ---- MPS (MPSImp.C, MPSImp_Variational.C) -----------------------------------------------
static void mpo_arrays(const MPO* o, std::vector<int>& dw1, std::vector<int>& dw2, std::vector<double>& w)
{
    for (int i = 1; i <= o->GetL(); i++) {
        const SiteOperator* so = o->GetSiteOperator(i);
        dw1.push_back(so->GetDw1());
        dw2.push_back(so->GetDw2());
        w.insert(w.end(), so->GetW().begin(), so->GetW().end());
    }
}

MPS::MPS(const Hamiltonian* H, int D, double normEps, double, int Dcap)
    : itsL(H->GetL()), itsd(H->Getd()), itsNormEps(normEps)
{
    itsCtx = nullptr;
    itsChain = nullptr;
    int rc = tn_ctx_create(0, &itsCtx);
    if (rc != 0) throw std::runtime_error("MPS: no HIP device (tn_ctx_create failed)");
    std::vector<int> dw1, dw2;
    std::vector<double> w;
    mpo_arrays(H, dw1, dw2, w);
    rc = tn_chain_create(itsCtx, itsL, itsd, dw1.data(), dw2.data(), w.data(), D, Dcap > 0 ? Dcap : D, &itsChain);
    if (rc < 0) {
        const std::string msg = std::string("tn_chain_create: ") + tn_last_error(itsCtx);
        tn_ctx_destroy(itsCtx);  // the destructor does not run for a failed constructor
        throw std::runtime_error(msg);
    }
    tn_sweep_opts_default(&itsOpts);
}

MPS::~MPS()
{
    if (itsChain) tn_chain_destroy(itsChain);
    if (itsCtx) tn_ctx_destroy(itsCtx);
}

void MPS::InitializeWith(State s, uint64_t seed)
{
    int st = s == Product ? 0 : (s == Random ? 1 : (s == Neel ? 2 : -1));
    if (st < 0) throw std::invalid_argument("InitializeWith: unsupported state");
    check(itsCtx, tn_chain_initialize(itsChain, st, seed), "tn_chain_initialize");
}

void MPS::Normalize(Direction LR) { check(itsCtx, tn_chain_normalize(itsChain, LR == DLeft ? 0 : 1), "tn_chain_normalize"); }

void MPS::NormalizeQR(Direction LR)
{
    check(itsCtx, tn_chain_normalize_qr(itsChain, LR == DLeft ? 0 : 1), "tn_chain_normalize_qr");
}

std::string MPS::GetNormStatus() const
{
    std::vector<char> buf(itsL + 1);
    check(itsCtx, tn_chain_norm_status(itsChain, itsNormEps, buf.data()), "tn_chain_norm_status");
    return std::string(buf.data());
}

char MPS::GetNormStatus(int isite) const
{
    if (isite < 1 || isite > itsL) throw std::invalid_argument("MPS::GetNormStatus: site out of range");
    return GetNormStatus()[isite - 1];
}

void MPS::Save(const std::string& path, bool withEnvironments) const
{
    check(itsCtx, tn_chain_save(itsChain, path.c_str(), nullptr, 0, withEnvironments ? 1 : 0), "tn_chain_save");
}

MPS* MPS::Load(const std::string& path, int Dcap)
{
    // L and d from the checkpoint header ("TNCHKPT1", version, L, d, ...)
    int head[5] = {0, 0, 0, 0, 0};
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::invalid_argument("MPS::Load: cannot open " + path);
    char magic[8];
    const bool ok = std::fread(magic, 1, 8, f) == 8 && std::fread(head, 4, 5, f) == 5;
    std::fclose(f);
    if (!ok) throw std::invalid_argument("MPS::Load: " + path + " is not a chain checkpoint");
    MPS* m = new MPS();
    if (tn_ctx_create(0, &m->itsCtx) != 0) {
        delete m;
        throw std::runtime_error("MPS::Load: no HIP device (tn_ctx_create failed)");
    }
    int64_t nb = 0;
    const int rc = tn_chain_load(m->itsCtx, path.c_str(), Dcap, &m->itsChain, nullptr, 0, &nb);
    if (rc < 0) {
        const std::string msg = std::string("tn_chain_load: ") + tn_last_error(m->itsCtx);
        delete m;
        throw std::runtime_error(msg);
    }
    m->itsL = head[1];
    m->itsd = head[2];
    tn_sweep_opts_default(&m->itsOpts);
    return m;
}

int MPS::GetMaxD() const
{
    std::vector<int> D1(itsL), D2(itsL);
    tn_chain_get_dims(itsChain, D1.data(), D2.data());
    return *std::max_element(D2.begin(), D2.end());
}

double MPS::GetExpectation(const MPO* o) const
{
    std::vector<int> dw1, dw2;
    std::vector<double> w;
    mpo_arrays(o, dw1, dw2, w);
    double v = 0;
    check(itsCtx, tn_chain_expectation(itsChain, dw1.data(), dw2.data(), w.data(), &v), "tn_chain_expectation");
    return v;
}

void MPS::Sweep(Direction lr, const Epsilons& eps)
{
    tn_sweep_opts o = itsOpts;
    o.eig_tol = std::min(o.eig_tol, eps.itsEigenSolverEpsilon);
    check(itsCtx, tn_chain_sweep(itsChain, lr == DLeft ? 0 : 1, &o, &itsReport.stats), "tn_chain_sweep");
}

double MPS::FindVariationalGroundState(const Hamiltonian* H, const IterationSchedule& is)
{
    Normalize(DRight);
    check(itsCtx, tn_chain_load_caches(itsChain, &itsOpts), "tn_chain_load_caches");
    double DE2 = 0;
    for (is.begin(); !is.end(); is++) DE2 = FindVariationalGroundState(H, *is);
    return DE2;
}

double MPS::FindVariationalGroundState(const Hamiltonian* H, const IterationScheduleLine& isl)
{
    const int Dmax = GetMaxD();
    MPO* H2 = H->CreateH2Operator();
    double DE2 = 0;
    for (int D = Dmax; D <= isl.itsDmax; D += isl.itsDeltaD) {
        if (D > Dmax) {
            check(itsCtx, tn_chain_increase_bond_dims(itsChain, D), "tn_chain_increase_bond_dims");
            Normalize(DRight);
            check(itsCtx, tn_chain_load_caches(itsChain, &itsOpts), "tn_chain_load_caches");
        }
        int in = 0;
        Logger->LogInfo(2, "iter    D   dE    dlambda");
        for (; in < isl.itsMaxGSSweepIterations; in++) {
            Sweep(DLeft, isl.itsEps);
            Sweep(DRight, isl.itsEps);
            double dE = 0, dl = 0;
            tn_chain_deltas(itsChain, &dE, &dl);
            Logger->LogInfoV(2, "%4d %4d %.1e %.1e", in, D, dE, dl);
            if (dE < isl.itsEps.itsDelatEnergy1Epsilon && dl < isl.itsEps.itsDeltaLambdaEpsilon) break;
        }
        in++;
        itsReport.nsweeps = in;
        const double E1 = GetExpectation(H);
        const double E2 = GetExpectation(H2);
        DE2 = E2 - E1 * E1;
        Logger->LogInfoV(0, "Variational GS D=%4d, %4d iterations, <E>=%.11f, <E^2>-<E>^2=%.2e", D, in,
                         E1 / std::max(itsL - 1, 1), DE2);
    }
    delete H2;
    return DE2;
}

// ---- Factory (FactoryImp.C:16-105) -----------------------------------------------------
Factory* Factory::GetFactory() { return new Factory(); }

Hamiltonian* Factory::Make1D_NN_HeisenbergHamiltonian(int L, double S, MPOForm f, double Jxy, double Jz, double hz)
{
    return new Hamiltonian(L, S, f, Jxy, Jz, hz, 0.0);
}
Hamiltonian* Factory::Make1D_NN_HeisenbergHamiltonian(int L, double S, MPOForm f, double Jxy, double Jz, double hz,
                                                      double Dz)
{
    return new Hamiltonian(L, S, f, Jxy, Jz, hz, Dz);
}
// Hamiltonian_1D_NN_TransverseIsing.C:26-46: H = J sum Sz Sz + hx sum Sx, Dw = 3
Hamiltonian* Factory::Make1D_NN_TransverseIsingHamiltonian(int L, double S, MPOForm f, double J, double hx)
{
    SpinOps o(S);
    BulkW w(3, o.d);
    if (f == RegularUpper) {
        w.set(0, 0, o.I); w.set(0, 1, o.Sz); w.set(0, 2, o.Sx, hx); w.set(1, 2, o.Sz, J); w.set(2, 2, o.I);
    } else {
        w.set(0, 0, o.I); w.set(1, 0, o.Sz); w.set(2, 0, o.Sx, hx); w.set(2, 1, o.Sz, J); w.set(2, 2, o.I);
    }
    return new Hamiltonian(L, S, f, 3, w.W);
}

// Hamiltonian_1D_2Body_LongRange.C:26-63 and Hamiltonian_1D_3Body_LongRange.C:26-63:
// Dw = 2 + NN(NN+1)/2, for each range i a string of i channels
static std::vector<double> long_range_W(const SpinOps& o, MPOForm f, double J, double hx, int NN, bool three)
{
    const int Dw = 2 + NN * (NN + 1) / 2;
    BulkW w(Dw, o.d);
    w.set(0, 0, o.I);
    w.set(Dw - 1, Dw - 1, o.I);
    const std::vector<double> Sz2 = o.mul(o.Sz, o.Sz);
    const std::vector<double>& open = three ? Sz2 : o.Sz;
    int ib = 0;
    for (int i = 1; i <= NN; i++) {
        for (int j = 1; j < i; j++) {
            const std::vector<double>& link = three ? o.Sz : o.I;
            const double c = three ? 1.0 / (j * j) : 1.0;
            if (f == RegularUpper) w.set(ib + j, ib + 1 + j, link, c);
            else w.set(ib + 1 + j, ib + j, link, c);
        }
        if (f == RegularUpper) {
            w.set(0, ib + 1, open);
            w.set(ib + i, Dw - 1, open, J / (i * i));
        } else {
            w.set(ib + 1, 0, open);
            w.set(Dw - 1, ib + i, open, J / (i * i));
        }
        ib += i;
    }
    if (f == RegularUpper) w.set(0, Dw - 1, o.Sx, hx);
    else w.set(Dw - 1, 0, o.Sx, hx);
    return w.W;
}

Hamiltonian* Factory::Make1D_2BodyLongRangeHamiltonian(int L, double S, MPOForm f, double J, double hx, int NN)
{
    if (NN < 1) throw std::invalid_argument("Make1D_2BodyLongRangeHamiltonian: NN >= 1");
    SpinOps o(S);
    return new Hamiltonian(L, S, f, 2 + NN * (NN + 1) / 2, long_range_W(o, f, J, hx, NN, false));
}

Hamiltonian* Factory::Make1D_3BodyLongRangeHamiltonian(int L, double S, MPOForm f, double J, double hx, int NN)
{
    if (NN < 1) throw std::invalid_argument("Make1D_3BodyLongRangeHamiltonian: NN >= 1");
    SpinOps o(S);
    return new Hamiltonian(L, S, f, 2 + NN * (NN + 1) / 2, long_range_W(o, f, J, hx, NN, true));
}

// Hamiltonian_1D_3Body.C:26-56 (entries as given there), Dw = 5
Hamiltonian* Factory::Make1D_3BodyHamiltonian(int L, double S, MPOForm f, double J, double K, double hz)
{
    SpinOps o(S);
    BulkW w(5, o.d);
    if (f == RegularUpper) {
        w.set(0, 0, o.I); w.set(0, 1, o.Sx); w.set(0, 2, o.Sx); w.set(0, 4, o.Sz, hz);
        w.set(3, 2, o.Sz); w.set(1, 4, o.Sx, J); w.set(3, 4, o.Sx, K); w.set(4, 4, o.I);
    } else {
        w.set(0, 0, o.I); w.set(1, 0, o.Sx); w.set(2, 0, o.Sx); w.set(4, 0, o.Sz, hz);
        w.set(2, 3, o.Sz); w.set(4, 1, o.Sx, J); w.set(4, 3, o.Sx, K); w.set(4, 4, o.I);
    }
    return new Hamiltonian(L, S, f, 5, w.W);
}

SVCompressorC* Factory::MakeMPSCompressor(int Dmax, double epsSV) { return new SVCompressorC(Dmax, epsSV); }
TNSLogger* Factory::MakeSPDLogger(int level) { return new TNSLogger(level); }

}  // namespace TensorNetworks
