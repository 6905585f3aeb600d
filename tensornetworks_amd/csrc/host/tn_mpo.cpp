// tn_mpo.cpp -- host-side MPO canonical form and compression (C ABI: tn_mpo_reduce).
//
// An MPO in regular upper form has, on every bond, a first channel that
// carries the identity from the left ("nothing placed yet") and a last channel
// that carries the identity to the right ("everything placed"); the channels
// in between (the interior) carry the interactions that cross the bond.  The
// reduction keeps both identity channels and shrinks the interior to its
// rank, following the block decomposition of D. E. Parker, X. Cao and
// M. P. Zaletel, Phys. Rev. B 102, 035147 (2020), the method behind the
// reference's MPO::CanonicalForm / MPO::Compress(Parker, Dmax, eps)
// (TensorNetworks/MPO.C:93-128, Operators/SiteOperatorImpCompress.C:16-100,
// Operators/OperatorValuedMatrix.C:437-537,786-840):
//
//   left step at a site W (n1 x n2 blocks of d x d):
//     V  = rows 0..n1-2 (row 0 only on the first site), columns 0..n2-2,
//          flattened to a (rows*d*d) x (n2-1) matrix;
//     V  = Q R (Householder), scaled so R[0][0] = 1 (column 0 of V is the
//          identity of norm sqrt(d));
//     W <- [Q | last column], the next site <- [[R, 0], [0, 1]] W_next.
//   compression adds, with R = [[1, r], [0, M]] = diag(1, M) [[1, r], [0, 1]]:
//     M  = U s V^T, truncated by the SVCompressor rule (SVCompressorImp.C:19-73),
//     W <- [Q diag(1, U) | last column],
//     next site <- [[1, 0, 0], [0, s V^T, 0], [0, 0, 1]] [[1, r, 0], [0, 1, 0], [0, 0, 1]] W_next.
// Right steps are left steps of the mirrored chain (site order reversed,
// channel matrix transposed and both channel orders reversed), which maps a
// regular upper MPO onto one.  A regular lower MPO is handled as the upper
// one with its channel order reversed on every bond.  Every step is an exact
// rewrite of the operator except the dropped singular values.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/tn_abi.h"

namespace {

struct Site {
    int n1 = 0, n2 = 0;
    std::vector<double> w;  // [(a*n2 + b)*d*d + m*d + n]
};

struct Chain {
    int d = 0;
    std::vector<Site> s;
    double* at(int i, int a, int b) { return &s[i].w[((size_t)a * s[i].n2 + b) * d * d]; }
};

// channel order reversed on both sides (lower <-> upper form)
void reverse_channels(Chain& c)
{
    const int dd = c.d * c.d;
    for (auto& st : c.s) {
        std::vector<double> w(st.w.size());
        for (int a = 0; a < st.n1; a++)
            for (int b = 0; b < st.n2; b++)
                std::memcpy(&w[((size_t)a * st.n2 + b) * dd], &st.w[((size_t)(st.n1 - 1 - a) * st.n2 + (st.n2 - 1 - b)) * dd],
                            dd * sizeof(double));
        st.w.swap(w);
    }
}

// W'[a][b] = W[n2-1-b][n1-1-a] and the site order reversed
void mirror(Chain& c)
{
    const int dd = c.d * c.d;
    std::reverse(c.s.begin(), c.s.end());
    for (auto& st : c.s) {
        std::vector<double> w(st.w.size());
        const int n1 = st.n2, n2 = st.n1;  // new shape
        for (int a = 0; a < n1; a++)
            for (int b = 0; b < n2; b++)
                std::memcpy(&w[((size_t)a * n2 + b) * dd], &st.w[((size_t)(st.n1 - 1 - b) * st.n2 + (st.n2 - 1 - a)) * dd],
                            dd * sizeof(double));
        st.w.swap(w);
        st.n1 = n1;
        st.n2 = n2;
    }
}

// Householder QR of a column-major m x n matrix: Q (m x k), R (k x n), k = min(m, n),
// diag(R) >= 0.
void householder_qr(int m, int n, std::vector<double> A, std::vector<double>& Q, std::vector<double>& R)
{
    const int k = std::min(m, n);
    std::vector<double> V((size_t)m * k, 0.0);
    for (int j = 0; j < k; j++) {
        double nx = 0;
        for (int i = j; i < m; i++) nx += A[i + (size_t)j * m] * A[i + (size_t)j * m];
        nx = std::sqrt(nx);
        if (nx == 0.0) continue;
        const double x0 = A[j + (size_t)j * m];
        const double alpha = x0 >= 0 ? -nx : nx;
        double* v = &V[(size_t)j * m];
        for (int i = j; i < m; i++) v[i] = A[i + (size_t)j * m];
        v[j] -= alpha;
        double nv = 0;
        for (int i = j; i < m; i++) nv += v[i] * v[i];
        nv = std::sqrt(nv);
        if (nv == 0.0) continue;
        for (int i = j; i < m; i++) v[i] /= nv;
        for (int c = j; c < n; c++) {
            double t = 0;
            for (int i = j; i < m; i++) t += v[i] * A[i + (size_t)c * m];
            for (int i = j; i < m; i++) A[i + (size_t)c * m] -= 2.0 * t * v[i];
        }
    }
    R.assign((size_t)k * n, 0.0);
    for (int i = 0; i < k; i++)
        for (int j = i; j < n; j++) R[i + (size_t)j * k] = A[i + (size_t)j * m];
    Q.assign((size_t)m * k, 0.0);
    for (int j = 0; j < k; j++) Q[j + (size_t)j * m] = 1.0;
    for (int j = k - 1; j >= 0; j--) {
        const double* v = &V[(size_t)j * m];
        for (int c = 0; c < k; c++) {
            double t = 0;
            for (int i = j; i < m; i++) t += v[i] * Q[i + (size_t)c * m];
            for (int i = j; i < m; i++) Q[i + (size_t)c * m] -= 2.0 * t * v[i];
        }
    }
    for (int i = 0; i < k; i++)
        if (R[i + (size_t)i * k] < 0.0) {
            for (int j = 0; j < n; j++) R[i + (size_t)j * k] = -R[i + (size_t)j * k];
            for (int r = 0; r < m; r++) Q[r + (size_t)i * m] = -Q[r + (size_t)i * m];
        }
}

// Thin SVD of a column-major p x q matrix by one-sided Jacobi:
// U (p x k), s (k, descending), Vt (k x q), k = min(p, q); U and Vt orthonormal.
void jacobi_svd(int p, int q, const std::vector<double>& M, std::vector<double>& U, std::vector<double>& s,
                std::vector<double>& Vt)
{
    const bool tr = p < q;
    const int r = tr ? q : p, c = tr ? p : q;  // work on X (r x c), r >= c
    std::vector<double> X((size_t)r * c), J((size_t)c * c, 0.0);
    for (int i = 0; i < r; i++)
        for (int j = 0; j < c; j++) X[i + (size_t)j * r] = tr ? M[j + (size_t)i * p] : M[i + (size_t)j * p];
    for (int j = 0; j < c; j++) J[j + (size_t)j * c] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        bool rotated = false;
        for (int a = 0; a < c - 1; a++)
            for (int b = a + 1; b < c; b++) {
                double al = 0, be = 0, ga = 0;
                for (int i = 0; i < r; i++) {
                    const double xa = X[i + (size_t)a * r], xb = X[i + (size_t)b * r];
                    al += xa * xa;
                    be += xb * xb;
                    ga += xa * xb;
                }
                if (ga == 0.0 || std::fabs(ga) <= 1e-15 * std::sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
                for (int i = 0; i < r; i++) {
                    const double xa = X[i + (size_t)a * r], xb = X[i + (size_t)b * r];
                    X[i + (size_t)a * r] = cs * xa - sn * xb;
                    X[i + (size_t)b * r] = sn * xa + cs * xb;
                }
                for (int i = 0; i < c; i++) {
                    const double ja = J[i + (size_t)a * c], jb = J[i + (size_t)b * c];
                    J[i + (size_t)a * c] = cs * ja - sn * jb;
                    J[i + (size_t)b * c] = sn * ja + cs * jb;
                }
            }
        if (!rotated) break;
    }
    std::vector<double> nrm(c);
    for (int j = 0; j < c; j++) {
        double t = 0;
        for (int i = 0; i < r; i++) t += X[i + (size_t)j * r] * X[i + (size_t)j * r];
        nrm[j] = std::sqrt(t);
    }
    std::vector<int> ord(c);
    for (int j = 0; j < c; j++) ord[j] = j;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nrm[x] > nrm[y]; });
    // left vectors of X: columns of X / s, completed to an orthonormal set where s = 0
    const double smax = c > 0 ? nrm[ord[0]] : 0.0;
    std::vector<double> Ux((size_t)r * c, 0.0);
    s.assign(c, 0.0);
    int e = 0;  // next unit vector tried for the completion
    for (int jj = 0; jj < c; jj++) {
        const int j = ord[jj];
        double* u = &Ux[(size_t)jj * r];
        if (nrm[j] > 1e-300 && nrm[j] > 1e-14 * smax) {
            s[jj] = nrm[j];
            for (int i = 0; i < r; i++) u[i] = X[i + (size_t)j * r] / nrm[j];
        } else {
            s[jj] = nrm[j];
            for (; e < r; e++) {
                std::fill(u, u + r, 0.0);
                u[e] = 1.0;
                for (int k = 0; k < jj; k++) {
                    const double* w = &Ux[(size_t)k * r];
                    double t = 0;
                    for (int i = 0; i < r; i++) t += w[i] * u[i];
                    for (int i = 0; i < r; i++) u[i] -= t * w[i];
                }
                double t = 0;
                for (int i = 0; i < r; i++) t += u[i] * u[i];
                if (t > 0.5) {
                    t = std::sqrt(t);
                    for (int i = 0; i < r; i++) u[i] /= t;
                    e++;
                    break;
                }
            }
        }
    }
    // M = X J^T (no transpose) or M^T = X J^T:  M = Ux s (J ord)^T  or  M = (J ord) s Ux^T
    const int k = c;
    U.assign((size_t)p * k, 0.0);
    Vt.assign((size_t)k * q, 0.0);
    for (int jj = 0; jj < k; jj++) {
        const int j = ord[jj];
        if (!tr) {
            for (int i = 0; i < p; i++) U[i + (size_t)jj * p] = Ux[i + (size_t)jj * r];
            for (int i = 0; i < q; i++) Vt[jj + (size_t)i * k] = J[i + (size_t)j * c];
        } else {
            for (int i = 0; i < p; i++) U[i + (size_t)jj * p] = J[i + (size_t)j * c];
            for (int i = 0; i < q; i++) Vt[jj + (size_t)i * k] = Ux[i + (size_t)jj * r];
        }
    }
}

// one left step at site i (0 <= i < L-1) of a regular upper chain; compress:
// also the SVD of the interior block with the SVCompressor rule (Dmax, eps).
// Returns the discarded weight.
double left_step(Chain& ch, int i, bool compress, int Dmax, double eps)
{
    const int d = ch.d, dd = d * d;
    Site& st = ch.s[i];
    Site& nx = ch.s[i + 1];
    const int n1 = st.n1, n2 = st.n2;
    if (n2 < 2) return 0.0;  // nothing but the identity channel crosses
    const int nr = n1 > 1 ? n1 - 1 : 1, nc = n2 - 1;
    const int m = nr * dd;
    // the last row must be [0 .. 0, I] (it is left out of V); tolerance for
    // the rounding of earlier steps, exact values are written back
    double scale = 0.0;
    for (double v : st.w) scale = std::max(scale, std::fabs(v));
    const double tol = 1e-12 * std::max(scale, 1.0);
    if (n1 > 1)
        for (int b = 0; b < nc; b++)
            for (int k = 0; k < dd; k++) {
                double* v = ch.at(i, n1 - 1, b) + k;
                if (std::fabs(*v) > tol) throw std::invalid_argument("MPO is not in regular form (last row)");
                *v = 0.0;
            }
    std::vector<double> V((size_t)m * nc);
    for (int r = 0; r < nr; r++)
        for (int b = 0; b < nc; b++)
            for (int k = 0; k < dd; k++) V[(size_t)r * dd + k + (size_t)b * m] = ch.at(i, r, b)[k];
    // column 0 must be the identity in row 0 only
    for (int r = 0; r < nr; r++)
        for (int a = 0; a < d; a++)
            for (int b = 0; b < d; b++) {
                double& v = V[(size_t)r * dd + a * d + b];
                const double want = (r == 0 && a == b) ? 1.0 : 0.0;
                if (std::fabs(v - want) > tol) throw std::invalid_argument("MPO is not in regular form (first column)");
                v = want;
            }
    std::vector<double> Q, R;
    householder_qr(m, nc, V, Q, R);
    const int k = std::min(m, nc);
    const double sd = std::sqrt((double)d);
    for (auto& q : Q) q *= sd;
    for (auto& x : R) x /= sd;
    // exactly: column 0 of V is the identity (norm sqrt(d)), so Q's column 0
    // is that identity and R's first column is e_0
    R[0] = 1.0;
    for (int r = 0; r < m; r++) Q[r] = r < dd && (r / d) == (r % d) ? 1.0 : 0.0;
    // Qx (m x kx), T (kx+1 x n2)
    std::vector<double> Qx, T;
    int kx = k;
    double disc = 0.0;
    std::vector<double> Tint;  // (kx-1) x (nc-1) interior rows of T
    if (compress && k > 1 && nc > 1) {
        const int p = k - 1, q = nc - 1;
        std::vector<double> M((size_t)p * q);
        for (int a = 0; a < p; a++)
            for (int b = 0; b < q; b++) M[a + (size_t)b * p] = R[(a + 1) + (size_t)(b + 1) * k];
        std::vector<double> U, s, Vt;
        jacobi_svd(p, q, M, U, s, Vt);
        const int ns = (int)s.size();
        int X = ns;
        double resc = 1.0;
        if (!(eps == 0.0 && Dmax <= 0)) {
            X = tn_svcompress_rank(ns, s.data(), Dmax, eps, &disc, &resc);
            if (X < 0) throw std::invalid_argument("compression rule failed");
        }
        kx = X + 1;
        Qx.assign((size_t)m * kx, 0.0);
        for (int r = 0; r < m; r++) Qx[r] = Q[r];
        for (int x = 0; x < X; x++)
            for (int a = 0; a < p; a++) {
                const double u = U[a + (size_t)x * p];
                if (u == 0.0) continue;
                for (int r = 0; r < m; r++) Qx[r + (size_t)(x + 1) * m] += Q[r + (size_t)(a + 1) * m] * u;
            }
        T.assign((size_t)(kx + 1) * n2, 0.0);
        // row 0: [1, r, 0]; rows 1..X: [0, s Vt, 0]; row kx: [0 .. 0, 1]
        T[0] = 1.0;
        for (int b = 1; b < nc; b++) T[0 + (size_t)b * (kx + 1)] = R[0 + (size_t)b * k];
        for (int x = 0; x < X; x++)
            for (int b = 1; b < nc; b++) T[(x + 1) + (size_t)b * (kx + 1)] = s[x] * resc * Vt[x + (size_t)(b - 1) * ns];
        T[kx + (size_t)(n2 - 1) * (kx + 1)] = 1.0;
    } else {
        Qx = Q;
        T.assign((size_t)(kx + 1) * n2, 0.0);
        for (int a = 0; a < k; a++)
            for (int b = 0; b < nc; b++) T[a + (size_t)b * (kx + 1)] = R[a + (size_t)b * k];
        T[kx + (size_t)(n2 - 1) * (kx + 1)] = 1.0;
    }
    // this site: [Qx | last column]
    Site ns;
    ns.n1 = n1;
    ns.n2 = kx + 1;
    ns.w.assign((size_t)n1 * ns.n2 * dd, 0.0);
    for (int r = 0; r < nr; r++)
        for (int x = 0; x < kx; x++)
            for (int e = 0; e < dd; e++) ns.w[((size_t)r * ns.n2 + x) * dd + e] = Qx[(size_t)r * dd + e + (size_t)x * m];
    for (int r = 0; r < n1; r++)
        std::memcpy(&ns.w[((size_t)r * ns.n2 + kx) * dd], ch.at(i, r, n2 - 1), dd * sizeof(double));
    // next site: T W_next
    Site nn;
    nn.n1 = kx + 1;
    nn.n2 = nx.n2;
    nn.w.assign((size_t)nn.n1 * nn.n2 * dd, 0.0);
    for (int a = 0; a <= kx; a++)
        for (int b = 0; b < n2; b++) {
            const double t = T[a + (size_t)b * (kx + 1)];
            if (t == 0.0) continue;
            for (int c = 0; c < nx.n2; c++) {
                const double* src = ch.at(i + 1, b, c);
                double* dst = &nn.w[((size_t)a * nn.n2 + c) * dd];
                for (int e = 0; e < dd; e++) dst[e] += t * src[e];
            }
        }
    st = ns;
    nx = nn;
    return disc;
}

double sweep_pair(Chain& ch, bool compress, int Dmax, double eps)
{
    const int L = (int)ch.s.size();
    double disc = 0.0;
    for (int i = 0; i + 1 < L; i++) disc += left_step(ch, i, compress, Dmax, eps);
    mirror(ch);
    for (int i = 0; i + 1 < L; i++) disc += left_step(ch, i, compress, Dmax, eps);
    mirror(ch);
    return disc;
}

}  // namespace

extern "C" int tn_mpo_reduce(int L, int d, const int* dw1, const int* dw2, const double* wcat, int form, int ops,
                             int Dmax, double eps, int* dw1_out, int* dw2_out, double* w_out, int64_t cap,
                             double* terror)
{
    if (L < 1 || d < 2 || !dw1 || !dw2 || !wcat || !dw1_out || !dw2_out || (form != 0 && form != 1) ||
        (ops & ~3) || ops == 0 || eps < 0.0 || cap < 0 || (cap > 0 && !w_out))
        return TN_ERR_ARG;
    try {
        Chain ch;
        ch.d = d;
        ch.s.resize(L);
        size_t off = 0;
        for (int i = 0; i < L; i++) {
            if (dw1[i] < 1 || dw2[i] < 1 || (i > 0 && dw1[i] != dw2[i - 1])) return TN_ERR_ARG;
            ch.s[i].n1 = dw1[i];
            ch.s[i].n2 = dw2[i];
            const size_t n = (size_t)dw1[i] * dw2[i] * d * d;
            ch.s[i].w.assign(wcat + off, wcat + off + n);
            off += n;
        }
        if (dw1[0] != 1 || dw2[L - 1] != 1) return TN_ERR_ARG;
        if (form == 1) reverse_channels(ch);
        double disc = 0.0;
        if (ops & 1) sweep_pair(ch, false, 0, 0.0);
        if (ops & 2) disc = sweep_pair(ch, true, Dmax, eps);
        if (form == 1) reverse_channels(ch);
        size_t total = 0;
        for (int i = 0; i < L; i++) total += ch.s[i].w.size();
        for (int i = 0; i < L; i++) {
            dw1_out[i] = ch.s[i].n1;
            dw2_out[i] = ch.s[i].n2;
        }
        if (terror) *terror = L > 1 ? std::sqrt(disc) / (L - 1) : 0.0;
        if ((int64_t)total > cap) return (int)std::min<size_t>(total, 0x7fffffff);  // size query / too small
        off = 0;
        for (int i = 0; i < L; i++) {
            std::memcpy(w_out + off, ch.s[i].w.data(), ch.s[i].w.size() * sizeof(double));
            off += ch.s[i].w.size();
        }
        return (int)total;
    } catch (const std::exception& e) {
        if (getenv("TN_MPO_DEBUG")) fprintf(stderr, "tn_mpo_reduce: %s\n", e.what());
        return TN_ERR_ARG;
    }
}
