// tn_kernels.hip -- CDNA4 (gfx950) kernels for the variational MPS sweep.
//
// * gemm_k     : LDS-tiled GEMM on the matrix cores (v_mfma 16x16, fp64/fp32/bf16
//                operands, fp32/fp64 accumulate), register-prefetch pipelined,
//                XCD-aware tile order.  The three stages of Heff*psi and of the
//                environment updates (MPSSite_Contractions.C:155-400) are GEMMs
//                once the MPO index is mixed by wmix_k.
// * wmix_k     : applies the sparse MPO tensor W[w][w'][m][n] between the GEMMs
//                (coalesced along the trailing bond index).
// * mdot/gemv  : Lanczos re-orthogonalisation (fp64, deterministic reductions).
// * jacobi_k   : one block pair of a one-sided block-Jacobi SVD (fp64).
#include "tn_kernels.h"

#include <type_traits>

namespace tn {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// scalar conversions
__device__ __forceinline__ bf16 f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float(((uint32_t)h) << 16); }

template <typename T> __device__ __forceinline__ T cvt_from_d(double v);
template <> __device__ __forceinline__ double cvt_from_d<double>(double v) { return v; }
template <> __device__ __forceinline__ float cvt_from_d<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ bf16 cvt_from_d<bf16>(double v) { return f2bf((float)v); }
template <typename T> __device__ __forceinline__ T cvt_from_f(float v);
template <> __device__ __forceinline__ double cvt_from_f<double>(float v) { return (double)v; }
template <> __device__ __forceinline__ float cvt_from_f<float>(float v) { return v; }
template <> __device__ __forceinline__ bf16 cvt_from_f<bf16>(float v) { return f2bf(v); }
__device__ __forceinline__ double to_d(double v) { return v; }
__device__ __forceinline__ double to_d(float v) { return (double)v; }
__device__ __forceinline__ double to_d(bf16 v) { return (double)bf2f(v); }
__device__ __forceinline__ float to_f(double v) { return (float)v; }
__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(bf16 v) { return bf2f(v); }
template <typename A, typename T> __device__ __forceinline__ A to_acc(T v)
{
    if constexpr (std::is_same<A, double>::value) return to_d(v);
    else return to_f(v);
}

// v = h0 + h1 + h2 + O(2^-24 |v|): each bfloat16 piece is the rounding of what
// the previous ones leave (the differences are exact in the wider type)
__device__ __forceinline__ void split3_val(double v, bf16& h0, bf16& h1, bf16& h2)
{
    h0 = f2bf((float)v);
    double r = v - (double)bf2f(h0);
    h1 = f2bf((float)r);
    r -= (double)bf2f(h1);
    h2 = f2bf((float)r);
}
__device__ __forceinline__ void split3_val(float v, bf16& h0, bf16& h1, bf16& h2)
{
    h0 = f2bf(v);
    float r = v - bf2f(h0);
    h1 = f2bf(r);
    r -= bf2f(h1);
    h2 = f2bf(r);
}

// ---------------------------------------------------------------------------
// MFMA traits: 16x16 output block per wave instruction
//   A frag: lane l holds A[l%16][kbase + (l/16)*KPL + 0..KPL-1]
//   B frag: lane l holds B[kbase + (l/16)*KPL + ..][l%16]
//   C    : lane l holds C[row(l, r)][l%16], r = 0..3, with
//          row = (l/16)*4 + r for the f32/bf16 instructions and
//          row = (l/16) + 4*r for f64 (measured on gfx950 with a one-wave probe)
template <typename TI> struct MT;
template <> struct MT<double> {
    static constexpr int KPL = 1, MK = 4, VEC = 2, PAD = 2;
    static __device__ __forceinline__ int crow(int lane, int r) { return (lane >> 4) + 4 * r; }
    typedef double frag_t;
    typedef v4d acc_t;
    typedef double acc_e;
};
template <> struct MT<float> {
    static constexpr int KPL = 1, MK = 4, VEC = 4, PAD = 4;
    static __device__ __forceinline__ int crow(int lane, int r) { return (lane >> 4) * 4 + r; }
    typedef float frag_t;
    typedef v4f acc_t;
    typedef float acc_e;
};
template <> struct MT<bf16> {
    static constexpr int KPL = 8, MK = 32, VEC = 8, PAD = 8;
    static __device__ __forceinline__ int crow(int lane, int r) { return (lane >> 4) * 4 + r; }
    typedef v8bf frag_t;
    typedef v4f acc_t;
    typedef float acc_e;
};

__device__ __forceinline__ v4d mma(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mma(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mma(v8bf a, v8bf b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

struct GemmParams {
    const void* A;
    const void* B;
    void* C;
    int M, N, K;
    long lda, ldb, ldc, sA, sB, sC;
    float alpha;
    int accumulate;
};

// element j of a 16-byte vector holding VEC values of type TI
template <typename TI> __device__ __forceinline__ TI vec_elem(const u32x4& v, int j);
template <> __device__ __forceinline__ double vec_elem<double>(const u32x4& v, int j)
{
    return j == 0 ? __hiloint2double((int)v.y, (int)v.x) : __hiloint2double((int)v.w, (int)v.z);
}
template <> __device__ __forceinline__ float vec_elem<float>(const u32x4& v, int j)
{
    const uint32_t w = j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
    return __uint_as_float(w);
}
template <> __device__ __forceinline__ bf16 vec_elem<bf16>(const u32x4& v, int j)
{
    const uint32_t w = (j >> 1) == 0 ? v.x : ((j >> 1) == 1 ? v.y : ((j >> 1) == 2 ? v.z : v.w));
    return (bf16)((j & 1) ? (w >> 16) : (w & 0xffffu));
}

// global -> registers for one operand tile (ROWS x BK, K-major in LDS)
//   KM  : operand(row, k) = G[row*ld + k]   (vector loads along k)
//   !KM : operand(row, k) = G[k*ld + row]   (vector loads along row)
template <typename TI, int ROWS, int BK, bool KM, int R>
__device__ __forceinline__ void tile_gload(u32x4 (&rg)[R], const TI* __restrict__ G, long ld,
                                           int r0, int k0, int tid)
{
    constexpr int VEC = MT<TI>::VEC;
    constexpr int TOT = ROWS * BK / VEC;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int v = tid + r * 256;
        if (TOT % 256 == 0 || v < TOT) {
            if constexpr (KM) {
                const int row = v / (BK / VEC), kv = v % (BK / VEC);
                rg[r] = *(const u32x4*)(G + (long)(r0 + row) * ld + k0 + kv * VEC);
            } else {
                const int kk = v / (ROWS / VEC), rv = v % (ROWS / VEC);
                rg[r] = *(const u32x4*)(G + (long)(k0 + kk) * ld + r0 + rv * VEC);
            }
        }
    }
}
template <typename TI, int ROWS, int BK, bool KM, int R>
__device__ __forceinline__ void tile_sstore(const u32x4 (&rg)[R], TI* __restrict__ S, int tid)
{
    constexpr int VEC = MT<TI>::VEC;
    constexpr int TOT = ROWS * BK / VEC;
    constexpr int LDK = BK + MT<TI>::PAD;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int v = tid + r * 256;
        if (TOT % 256 == 0 || v < TOT) {
            if constexpr (KM) {
                const int row = v / (BK / VEC), kv = v % (BK / VEC);
                *(u32x4*)(S + row * LDK + kv * VEC) = rg[r];
            } else {
                const int kk = v / (ROWS / VEC), rv = v % (ROWS / VEC);
#pragma unroll
                for (int j = 0; j < VEC; j++) S[(rv * VEC + j) * LDK + kk] = vec_elem<TI>(rg[r], j);
            }
        }
    }
}

// One BM x BN output tile: 256 threads = 4 waves in a 2x2 arrangement, wave
// tile (BM/2) x (BN/2); operand tiles staged through LDS (As, Bs) with the
// next K-slice prefetched into registers.
template <typename TI, typename TO, int BM, int BN, int BK, bool AKM, bool BKM>
__device__ __forceinline__ void gemm_tile(const TI* __restrict__ A, const TI* __restrict__ B, TO* __restrict__ C,
                                          long lda, long ldb, long ldc, int K, float alpha, int accumulate,
                                          int m0, int n0, TI* __restrict__ As, TI* __restrict__ Bs)
{
    typedef MT<TI> T;
    typedef typename T::frag_t frag_t;
    typedef typename T::acc_t acc_t;
    constexpr int LDK = BK + T::PAD;
    constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
    constexpr int VEC = T::VEC;
    constexpr int TA = BM * BK / VEC, TB = BN * BK / VEC;
    constexpr int RA = (TA + 255) / 256, RB = (TB + 255) / 256;
    static_assert(FM >= 1 && FN >= 1, "tile");
    static_assert(BK % T::MK == 0, "BK");

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm0 = (wid >> 1) * WM, wn0 = (wid & 1) * WN;

    u32x4 ra[RA], rb[RB];
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[i][j][r] = 0;

    const int nk = K / BK;
    tile_gload<TI, BM, BK, AKM>(ra, A, lda, m0, 0, tid);
    tile_gload<TI, BN, BK, BKM>(rb, B, ldb, n0, 0, tid);
    tile_sstore<TI, BM, BK, AKM>(ra, As, tid);
    tile_sstore<TI, BN, BK, BKM>(rb, Bs, tid);
    __syncthreads();
    for (int t = 0; t < nk; t++) {
        if (t + 1 < nk) {
            tile_gload<TI, BM, BK, AKM>(ra, A, lda, m0, (t + 1) * BK, tid);
            tile_gload<TI, BN, BK, BKM>(rb, B, ldb, n0, (t + 1) * BK, tid);
        }
        // the MFMA section of a slice at raised wave priority: a wave that has
        // its operands in LDS issues ahead of waves still loading or storing
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int kk = 0; kk < BK; kk += T::MK) {
            frag_t a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; i++)
                a[i] = *(const frag_t*)(As + (wm0 + i * 16 + (lane & 15)) * LDK + kk + (lane >> 4) * T::KPL);
#pragma unroll
            for (int j = 0; j < FN; j++)
                b[j] = *(const frag_t*)(Bs + (wn0 + j * 16 + (lane & 15)) * LDK + kk + (lane >> 4) * T::KPL);
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < FN; j++) acc[i][j] = mma(a[i], b[j], acc[i][j]);
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (t + 1 < nk) {
            tile_sstore<TI, BM, BK, AKM>(ra, As, tid);
            tile_sstore<TI, BN, BK, BKM>(rb, Bs, tid);
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long row = m0 + wm0 + i * 16 + T::crow(lane, r);
                const long col = n0 + wn0 + j * 16 + (lane & 15);
                TO* cp = C + row * ldc + col;
                if constexpr (std::is_same<TO, double>::value) {
                    double v = (double)acc[i][j][r] * (double)alpha;
                    if (accumulate) v += to_d(*cp);
                    *cp = cvt_from_d<TO>(v);
                } else {
                    float v = (float)acc[i][j][r] * alpha;
                    if (accumulate) v += to_f(*cp);
                    *cp = cvt_from_f<TO>(v);
                }
            }
}

template <typename TI, typename TO, int BM, int BN, int BK, bool AKM, bool BKM>
__global__ void __launch_bounds__(256) gemm_k(GemmParams p)
{
    constexpr int LDK = BK + MT<TI>::PAD;
    __shared__ __attribute__((aligned(16))) TI As[BM * LDK];
    __shared__ __attribute__((aligned(16))) TI Bs[BN * LDK];
    const int ntn = p.N / BN, ntm = p.M / BM;
    const int nblk = ntn * ntm;
    int bid = blockIdx.x;
    // consecutive block ids land on different XCDs; give each XCD a contiguous
    // range of output tiles so tiles sharing operands share an L2
    if ((nblk & 7) == 0) bid = (bid & 7) * (nblk >> 3) + (bid >> 3);
    // inside the range, GM tile rows at a time and column by column within
    // them: the tiles resident together on an XCD share GM row blocks of A and
    // a few column blocks of B, instead of one row block of A and all of B
    constexpr int GM = 8;
    const int gsz = GM * ntn, g = bid / gsz, first = g * GM;
    const int rows = ntm - first < GM ? ntm - first : GM;
    const int local = bid - g * gsz;
    const int tm = first + local % rows, tn = local / rows;
    const long bz = blockIdx.y;
    gemm_tile<TI, TO, BM, BN, BK, AKM, BKM>((const TI*)p.A + bz * p.sA, (const TI*)p.B + bz * p.sB,
                                            (TO*)p.C + bz * p.sC, p.lda, p.ldb, p.ldc, p.K, p.alpha,
                                            p.accumulate, tm * BM, tn * BN, As, Bs);
}

// Split-product LDS layout for BK = 32: rows of 64 bytes without padding,
// the four 16-byte k-units of row r stored in the order u ^ g(r), g(r) =
// ((r >> 1) ^ (r >> 2)) & 3.  A tile store (lanes 4i .. 4i+3 write the four
// units of one row) and a fragment read (lanes 0..15 read one unit of rows
// 0..15) then fall on distinct banks in every group of 8 or 16 lanes, with
// 32 or with 64 banks of 4 bytes; the padded layout (80-byte rows) leaves the
// stores 2-way conflicted.
__device__ __forceinline__ int x3_swz(int row, int unit) { return row * 32 + ((unit ^ ((row >> 1) ^ (row >> 2))) & 3) * 8; }

template <int ROWS, int BK, int R>
__device__ __forceinline__ void x3_sstore(const u32x4 (&rg)[R], bf16* __restrict__ S, int tid)
{
    constexpr int UPR = BK / 8;  // 16-byte units per row
    constexpr int TOT = ROWS * UPR;
    constexpr int LDK = BK == 32 ? BK : BK + MT<bf16>::PAD;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int v = tid + r * 256;
        if (TOT % 256 == 0 || v < TOT) {
            const int row = v / UPR, kv = v % UPR;
            if constexpr (BK == 32) *(u32x4*)(S + x3_swz(row, kv)) = rg[r];
            else *(u32x4*)(S + row * LDK + kv * 8) = rg[r];
        }
    }
}

// One BM x BN output tile of the split-bfloat16 product: A and B each held as
// three bfloat16 planes (A = A0 + A1 + A2), K-major, every plane staged
// through LDS; each K step issues the six leading cross products A_i B_j
// (i + j <= 2) of v_mfma_f32_16x16x32_bf16 into one FP32 accumulator.  Same
// wave layout, register prefetch and wave-priority scheme as gemm_tile.
template <typename TO, int BM, int BN, int BK>
__device__ __forceinline__ void gemm_tile_x3(const bf16* __restrict__ A, long psA, const bf16* __restrict__ B,
                                             long psB, TO* __restrict__ C, long lda, long ldb, long ldc, int K,
                                             float alpha, int accumulate, int m0, int n0, bf16* __restrict__ As,
                                             bf16* __restrict__ Bs)
{
    typedef MT<bf16> T;
    typedef T::frag_t frag_t;
    typedef T::acc_t acc_t;
    constexpr bool SWZ = BK == 32;
    constexpr int LDK = SWZ ? BK : BK + T::PAD;
    constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
    constexpr int TA = BM * BK / T::VEC, TB = BN * BK / T::VEC;
    constexpr int RA = (TA + 255) / 256, RB = (TB + 255) / 256;
    constexpr int SA = BM * LDK, SB = BN * LDK;
    static_assert(FM >= 1 && FN >= 1 && BK % T::MK == 0, "tile");

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm0 = (wid >> 1) * WM, wn0 = (wid & 1) * WN;
    // LDS element offset of the fragment of row `row` at k-step kk
    auto foff = [&](int row, int kk) {
        if constexpr (SWZ) return x3_swz(row, (lane >> 4) + kk / 8);
        else return row * LDK + kk + (lane >> 4) * T::KPL;
    };

    u32x4 ra[3][RA], rb[3][RB];
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[i][j][r] = 0;

    const int nk = K / BK;
#pragma unroll
    for (int p = 0; p < 3; p++) {
        tile_gload<bf16, BM, BK, true>(ra[p], A + p * psA, lda, m0, 0, tid);
        tile_gload<bf16, BN, BK, true>(rb[p], B + p * psB, ldb, n0, 0, tid);
    }
#pragma unroll
    for (int p = 0; p < 3; p++) {
        x3_sstore<BM, BK>(ra[p], As + p * SA, tid);
        x3_sstore<BN, BK>(rb[p], Bs + p * SB, tid);
    }
    __syncthreads();
    for (int t = 0; t < nk; t++) {
        if (t + 1 < nk) {
#pragma unroll
            for (int p = 0; p < 3; p++) {
                tile_gload<bf16, BM, BK, true>(ra[p], A + p * psA, lda, m0, (t + 1) * BK, tid);
                tile_gload<bf16, BN, BK, true>(rb[p], B + p * psB, ldb, n0, (t + 1) * BK, tid);
            }
        }
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int kk = 0; kk < BK; kk += T::MK) {
            frag_t a[3][FM], b[3][FN];
#pragma unroll
            for (int p = 0; p < 3; p++) {
#pragma unroll
                for (int i = 0; i < FM; i++)
                    a[p][i] = *(const frag_t*)(As + p * SA + foff(wm0 + i * 16 + (lane & 15), kk));
#pragma unroll
                for (int j = 0; j < FN; j++)
                    b[p][j] = *(const frag_t*)(Bs + p * SB + foff(wn0 + j * 16 + (lane & 15), kk));
            }
#pragma unroll
            for (int i = 0; i < FM; i++)
#pragma unroll
                for (int j = 0; j < FN; j++) {
                    // smallest terms first
                    acc[i][j] = mma(a[2][i], b[0][j], acc[i][j]);
                    acc[i][j] = mma(a[0][i], b[2][j], acc[i][j]);
                    acc[i][j] = mma(a[1][i], b[1][j], acc[i][j]);
                    acc[i][j] = mma(a[1][i], b[0][j], acc[i][j]);
                    acc[i][j] = mma(a[0][i], b[1][j], acc[i][j]);
                    acc[i][j] = mma(a[0][i], b[0][j], acc[i][j]);
                }
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (t + 1 < nk) {
#pragma unroll
            for (int p = 0; p < 3; p++) {
                x3_sstore<BM, BK>(ra[p], As + p * SA, tid);
                x3_sstore<BN, BK>(rb[p], Bs + p * SB, tid);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long row = m0 + wm0 + i * 16 + T::crow(lane, r);
                const long col = n0 + wn0 + j * 16 + (lane & 15);
                TO* cp = C + row * ldc + col;
                if constexpr (std::is_same<TO, double>::value) {
                    double v = (double)acc[i][j][r] * (double)alpha;
                    if (accumulate) v += *cp;
                    *cp = v;
                } else {
                    float v = acc[i][j][r] * alpha;
                    if (accumulate) v += to_f(*cp);
                    *cp = cvt_from_f<TO>(v);
                }
            }
}

struct GemmX3Params {
    const bf16* A;
    const bf16* B;
    void* C;
    int M, N, K;
    long lda, ldb, ldc, sA, sB, sC, psA, psB;
    float alpha;
    int accumulate;
};

template <typename TO, int BM, int BN, int BK>
__global__ void __launch_bounds__(256) gemm_x3_k(GemmX3Params p)
{
    constexpr int LDK = BK == 32 ? BK : BK + MT<bf16>::PAD;
    __shared__ __attribute__((aligned(16))) bf16 As[3 * BM * LDK];
    __shared__ __attribute__((aligned(16))) bf16 Bs[3 * BN * LDK];
    const int ntn = p.N / BN, ntm = p.M / BM;
    const int nblk = ntn * ntm;
    int bid = blockIdx.x;
    // XCD-aware tile order as in gemm_k
    if ((nblk & 7) == 0) bid = (bid & 7) * (nblk >> 3) + (bid >> 3);
    constexpr int GM = 8;
    const int gsz = GM * ntn, g = bid / gsz, first = g * GM;
    const int rows = ntm - first < GM ? ntm - first : GM;
    const int local = bid - g * gsz;
    const int tm = first + local % rows, tn = local / rows;
    const long bz = blockIdx.y;
    gemm_tile_x3<TO, BM, BN, BK>(p.A + bz * p.sA, p.psA, p.B + bz * p.sB, p.psB, (TO*)p.C + bz * p.sC, p.lda, p.ldb,
                                 p.ldc, p.K, p.alpha, p.accumulate, tm * BM, tn * BN, As, Bs);
}

// The split product with eight waves per workgroup and one workgroup per CU
// (the whole chip in one round at the bulk shapes: 3072 x 4096 in 192 x 256
// tiles, 1024 x 3072 in 256 x 192 tiles with a 4-way K split).  Wave grid
// WGM x WGN, wave tile (BM / WGM) x (BN / WGN) = 96 x 64 or 64 x 96: 0.28 KB
// of LDS traffic per MFMA against 0.44 KB for the 96 x 128 four-wave tile.
// Same swizzled LDS layout, register prefetch and priority scheme; the
// fragments of B are loaded column block by column block to bound registers.
template <int ROWS, int R>
__device__ __forceinline__ void x3w_gload(u32x4 (&rg)[R], const bf16* __restrict__ G, long ld, int r0, int k0, int tid)
{
    constexpr int TOT = ROWS * 4;  // 16-byte units of a 32-wide K slice
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int v = tid + r * 512;
        if (TOT % 512 == 0 || v < TOT) rg[r] = *(const u32x4*)(G + (long)(r0 + v / 4) * ld + k0 + (v % 4) * 8);
    }
}
template <int ROWS, int R>
__device__ __forceinline__ void x3w_sstore(const u32x4 (&rg)[R], bf16* __restrict__ S, int tid)
{
    constexpr int TOT = ROWS * 4;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int v = tid + r * 512;
        if (TOT % 512 == 0 || v < TOT) *(u32x4*)(S + x3_swz(v / 4, v % 4)) = rg[r];
    }
}

template <typename TO, int BM, int BN, int WGM, int WGN, bool DB>
__global__ void __launch_bounds__(512) gemm_x3w_k(GemmX3Params p)
{
    typedef MT<bf16> T;
    typedef T::frag_t frag_t;
    typedef T::acc_t acc_t;
    static_assert(WGM * WGN == 8, "eight waves");
    constexpr int BK = 32;
    constexpr int WM = BM / WGM, WN = BN / WGN, FM = WM / 16, FN = WN / 16;
    static_assert(FM * 16 == WM && FN * 16 == WN, "wave tile");
    constexpr int RA = (BM * 4 + 511) / 512, RB = (BN * 4 + 511) / 512;
    constexpr int SA = BM * BK, SB = BN * BK;
    // DB: two LDS stages, the next K slice stored into the other stage while
    // the current one is in use, one barrier per step
    constexpr int NS = DB ? 2 : 1;
    __shared__ __attribute__((aligned(16))) bf16 As[NS * 3 * SA];
    __shared__ __attribute__((aligned(16))) bf16 Bs[NS * 3 * SB];

    const int ntn = p.N / BN, ntm = p.M / BM;
    const int nblk = ntn * ntm;
    int bid = blockIdx.x;
    if ((nblk & 7) == 0) bid = (bid & 7) * (nblk >> 3) + (bid >> 3);
    constexpr int GM = 8;
    const int gsz = GM * ntn, g = bid / gsz, first = g * GM;
    const int rows = ntm - first < GM ? ntm - first : GM;
    const int local = bid - g * gsz;
    const int m0 = (first + local % rows) * BM, n0 = (local / rows) * BN;
    const long bz = blockIdx.y;
    const bf16* __restrict__ A = p.A + bz * p.sA;
    const bf16* __restrict__ B = p.B + bz * p.sB;
    TO* __restrict__ C = (TO*)p.C + bz * p.sC;
    const long lda = p.lda, ldb = p.ldb, psA = p.psA, psB = p.psB;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm0 = (wid / WGN) * WM, wn0 = (wid % WGN) * WN;

    u32x4 ra[3][RA], rb[3][RB];
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[i][j][r] = 0;

    const int nk = p.K / BK;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        x3w_gload<BM>(ra[q], A + q * psA, lda, m0, 0, tid);
        x3w_gload<BN>(rb[q], B + q * psB, ldb, n0, 0, tid);
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
        x3w_sstore<BM>(ra[q], As + q * SA, tid);
        x3w_sstore<BN>(rb[q], Bs + q * SB, tid);
    }
    __syncthreads();
    for (int t = 0; t < nk; t++) {
        const int st = DB ? (t & 1) : 0;
        const bf16* Ac = As + st * 3 * SA;
        const bf16* Bc = Bs + st * 3 * SB;
        if (t + 1 < nk) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
                x3w_gload<BM>(ra[q], A + q * psA, lda, m0, (t + 1) * BK, tid);
                x3w_gload<BN>(rb[q], B + q * psB, ldb, n0, (t + 1) * BK, tid);
            }
        }
        __builtin_amdgcn_s_setprio(1);
        frag_t a[3][FM];
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int i = 0; i < FM; i++)
                a[q][i] = *(const frag_t*)(Ac + q * SA + x3_swz(wm0 + i * 16 + (lane & 15), lane >> 4));
#pragma unroll
        for (int j = 0; j < FN; j++) {
            frag_t b[3];
#pragma unroll
            for (int q = 0; q < 3; q++)
                b[q] = *(const frag_t*)(Bc + q * SB + x3_swz(wn0 + j * 16 + (lane & 15), lane >> 4));
#pragma unroll
            for (int i = 0; i < FM; i++) {
                // smallest terms first
                acc[i][j] = mma(a[2][i], b[0], acc[i][j]);
                acc[i][j] = mma(a[0][i], b[2], acc[i][j]);
                acc[i][j] = mma(a[1][i], b[1], acc[i][j]);
                acc[i][j] = mma(a[1][i], b[0], acc[i][j]);
                acc[i][j] = mma(a[0][i], b[1], acc[i][j]);
                acc[i][j] = mma(a[0][i], b[0], acc[i][j]);
            }
        }
        __builtin_amdgcn_s_setprio(0);
        if constexpr (DB) {
            // the other stage was last read in step t - 1, before the barrier
            // that ended it
            if (t + 1 < nk) {
                bf16* An = As + (st ^ 1) * 3 * SA;
                bf16* Bn = Bs + (st ^ 1) * 3 * SB;
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    x3w_sstore<BM>(ra[q], An + q * SA, tid);
                    x3w_sstore<BN>(rb[q], Bn + q * SB, tid);
                }
            }
            __syncthreads();
        } else {
            __syncthreads();
            if (t + 1 < nk) {
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    x3w_sstore<BM>(ra[q], As + q * SA, tid);
                    x3w_sstore<BN>(rb[q], Bs + q * SB, tid);
                }
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const long row = m0 + wm0 + i * 16 + T::crow(lane, r);
                const long col = n0 + wn0 + j * 16 + (lane & 15);
                TO* cp = C + row * p.ldc + col;
                if constexpr (std::is_same<TO, double>::value) {
                    double v = (double)acc[i][j][r] * (double)p.alpha;
                    if (p.accumulate) v += *cp;
                    *cp = v;
                } else {
                    float v = acc[i][j][r] * p.alpha;
                    if (p.accumulate) v += to_f(*cp);
                    *cp = cvt_from_f<TO>(v);
                }
            }
}

// Grouped GEMM: one persistent launch over the output tiles of many problems
// (problem q owns tiles [off[q], off[q+1]), row-major over its N/BN columns).
// Workgroups stride over the tile list; the owning problem is found by a
// binary search of off.
template <typename TI, typename TO, int BM, int BN, int BK, bool AKM, bool BKM>
__global__ void __launch_bounds__(256) gemm_grouped_k(const GemmProb* __restrict__ probs,
                                                      const int* __restrict__ off, int nprob)
{
    constexpr int LDK = BK + MT<TI>::PAD;
    __shared__ __attribute__((aligned(16))) TI As[BM * LDK];
    __shared__ __attribute__((aligned(16))) TI Bs[BN * LDK];
    const int total = off[nprob];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const GemmProb& q = probs[lo];
        const int local = t - off[lo], ntn = q.N / BN;
        const int tm = local / ntn, tn = local - tm * ntn;
        __syncthreads();  // LDS of the previous tile is free
        gemm_tile<TI, TO, BM, BN, BK, AKM, BKM>((const TI*)q.A, (const TI*)q.B, (TO*)q.C, q.lda, q.ldb, q.ldc, q.K,
                                                q.alpha, q.accumulate, tm * BM, tn * BN, As, Bs);
    }
}

// tile menu (code -> BM x BN; "deep" tiles stage twice the K slice per LDS
// round, halving the barriers per flop at twice the LDS footprint).  The
// 96-wide tiles exist for the split-bfloat16 product only: at two
// workgroups per CU, 3072 x 4096 in 96 x 128 tiles (and 1024 x 3072 in
// 128 x 96 tiles, split-K 2) fills the 256 CUs in whole rounds, where
// 128 x 128 tiles leave the last round half empty.
struct TileCfg { int bm, bn; bool deep; const char* name; bool x3only; };
static const TileCfg kTiles[] = {
    {128, 128, false, "128x128", false}, {128, 64, false, "128x64", false}, {64, 128, false, "64x128", false},
    {64, 64, false, "64x64", false},     {32, 64, false, "32x64", false},   {64, 32, false, "64x32", false},
    {32, 32, false, "32x32", false},     {128, 128, true, "128x128d", false}, {128, 64, true, "128x64d", false},
    {64, 128, true, "64x128d", false},   {96, 128, false, "96x128", true},  {128, 96, false, "128x96", true},
    {192, 256, false, "192x256w8", true}, {256, 192, false, "256x192w8", true},
    {192, 128, false, "192x128w8db", true}, {128, 192, false, "128x192w8db", true},
};
static const int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);
int gemm_num_tiles() { return kNumTiles; }
const char* gemm_tile_name(int code) { return (code >= 1 && code <= kNumTiles) ? kTiles[code - 1].name : "auto"; }

template <typename TI, typename TO, int BK, bool AKM, bool BKM>
static hipError_t launch_tile(int code, const GemmParams& p, int batch, hipStream_t s)
{
    const TileCfg& t = kTiles[code];
    dim3 grid((p.M / t.bm) * (p.N / t.bn), batch);
    switch (code) {
    case 0: gemm_k<TI, TO, 128, 128, BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    case 1: gemm_k<TI, TO, 128, 64, BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    case 2: gemm_k<TI, TO, 64, 128, BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    case 3: gemm_k<TI, TO, 64, 64, BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    case 4: gemm_k<TI, TO, 32, 64, BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    case 5: gemm_k<TI, TO, 64, 32, BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    case 6: gemm_k<TI, TO, 32, 32, BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    case 7: gemm_k<TI, TO, 128, 128, 2 * BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    case 8: gemm_k<TI, TO, 128, 64, 2 * BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    default: gemm_k<TI, TO, 64, 128, 2 * BK, AKM, BKM><<<grid, 256, 0, s>>>(p); break;
    }
    return hipGetLastError();
}

template <typename TI, typename TO, int BK, bool AKM, bool BKM>
static hipError_t launch_grouped(int code, const GemmProb* probs, const int* off, int nprob, int grid,
                                 hipStream_t s)
{
    switch (code) {
    case 0: gemm_grouped_k<TI, TO, 128, 128, BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    case 1: gemm_grouped_k<TI, TO, 128, 64, BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    case 2: gemm_grouped_k<TI, TO, 64, 128, BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    case 3: gemm_grouped_k<TI, TO, 64, 64, BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    case 4: gemm_grouped_k<TI, TO, 32, 64, BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    case 5: gemm_grouped_k<TI, TO, 64, 32, BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    case 6: gemm_grouped_k<TI, TO, 32, 32, BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    case 7: gemm_grouped_k<TI, TO, 128, 128, 2 * BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    case 8: gemm_grouped_k<TI, TO, 128, 64, 2 * BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    default: gemm_grouped_k<TI, TO, 64, 128, 2 * BK, AKM, BKM><<<grid, 256, 0, s>>>(probs, off, nprob); break;
    }
    return hipGetLastError();
}

template <typename TI, typename TO, bool AKM, bool BKM>
static hipError_t grouped_typed(int code, int bk, const GemmProb* probs, const int* off, int nprob, int grid,
                                hipStream_t s)
{
    if constexpr (std::is_same<TI, double>::value) {
        return launch_grouped<TI, TO, 16, AKM, BKM>(code, probs, off, nprob, grid, s);
    } else if constexpr (std::is_same<TI, float>::value) {
        if (bk == 32) return launch_grouped<TI, TO, 32, AKM, BKM>(code, probs, off, nprob, grid, s);
        return launch_grouped<TI, TO, 16, AKM, BKM>(code, probs, off, nprob, grid, s);
    } else {
        if (bk == 64) return launch_grouped<TI, TO, 64, AKM, BKM>(code, probs, off, nprob, grid, s);
        return launch_grouped<TI, TO, 32, AKM, BKM>(code, probs, off, nprob, grid, s);
    }
}

template <bool AKM, bool BKM>
static hipError_t grouped_layout(int tin, int tout, int code, int bk, const GemmProb* probs, const int* off,
                                 int nprob, int grid, hipStream_t s)
{
    if (tin == T_F64 && tout == T_F64) return grouped_typed<double, double, AKM, BKM>(code, bk, probs, off, nprob, grid, s);
    if (tin == T_F32 && tout == T_F32) return grouped_typed<float, float, AKM, BKM>(code, bk, probs, off, nprob, grid, s);
    if (tin == T_F32 && tout == T_F64) return grouped_typed<float, double, AKM, BKM>(code, bk, probs, off, nprob, grid, s);
    if (tin == T_BF16 && tout == T_F32) return grouped_typed<bf16, float, AKM, BKM>(code, bk, probs, off, nprob, grid, s);
    if (tin == T_BF16 && tout == T_F64) return grouped_typed<bf16, double, AKM, BKM>(code, bk, probs, off, nprob, grid, s);
    return hipErrorInvalidValue;
}

template <typename TO>
static hipError_t launch_x3(int code, const GemmX3Params& p, int batch, hipStream_t s)
{
    const TileCfg& t = kTiles[code];
    dim3 grid((p.M / t.bm) * (p.N / t.bn), batch);
    switch (code) {
    case 0: gemm_x3_k<TO, 128, 128, 32><<<grid, 256, 0, s>>>(p); break;
    case 1: gemm_x3_k<TO, 128, 64, 32><<<grid, 256, 0, s>>>(p); break;
    case 2: gemm_x3_k<TO, 64, 128, 32><<<grid, 256, 0, s>>>(p); break;
    case 3: gemm_x3_k<TO, 64, 64, 32><<<grid, 256, 0, s>>>(p); break;
    case 4: gemm_x3_k<TO, 32, 64, 32><<<grid, 256, 0, s>>>(p); break;
    case 5: gemm_x3_k<TO, 64, 32, 32><<<grid, 256, 0, s>>>(p); break;
    case 6: gemm_x3_k<TO, 32, 32, 32><<<grid, 256, 0, s>>>(p); break;
    // deep tiles: a K step of 64 per LDS round (half the barriers per MFMA)
    case 7: gemm_x3_k<TO, 128, 128, 64><<<grid, 256, 0, s>>>(p); break;
    case 8: gemm_x3_k<TO, 128, 64, 64><<<grid, 256, 0, s>>>(p); break;
    case 9: gemm_x3_k<TO, 64, 128, 64><<<grid, 256, 0, s>>>(p); break;
    case 10: gemm_x3_k<TO, 96, 128, 32><<<grid, 256, 0, s>>>(p); break;
    case 11: gemm_x3_k<TO, 128, 96, 32><<<grid, 256, 0, s>>>(p); break;
    // eight waves, one workgroup per CU
    case 12: gemm_x3w_k<TO, 192, 256, 2, 4, false><<<grid, 512, 0, s>>>(p); break;
    case 13: gemm_x3w_k<TO, 256, 192, 4, 2, false><<<grid, 512, 0, s>>>(p); break;
    // eight waves, two LDS stages
    case 14: gemm_x3w_k<TO, 192, 128, 4, 2, true><<<grid, 512, 0, s>>>(p); break;
    default: gemm_x3w_k<TO, 128, 192, 2, 4, true><<<grid, 512, 0, s>>>(p); break;
    }
    return hipGetLastError();
}

static int choose_tile(int M, int N, int batch);
// the split-bfloat16 product: K-major planes only, K a multiple of 32
static hipError_t gemm_x3(const GemmDesc& g, hipStream_t s)
{
    if (!g.a_kmajor || !g.b_kmajor || g.K % 32 || g.psA <= 0 || g.psB <= 0) return hipErrorInvalidValue;
    GemmX3Params p;
    p.A = (const bf16*)g.A; p.B = (const bf16*)g.B; p.C = g.C;
    p.M = g.M; p.N = g.N; p.K = g.K;
    p.lda = g.lda; p.ldb = g.ldb; p.ldc = g.ldc;
    p.sA = g.sA; p.sB = g.sB; p.sC = g.sC;
    p.psA = g.psA; p.psB = g.psB;
    p.alpha = g.alpha; p.accumulate = g.accumulate;
    const int code = g.tile > 0 ? g.tile - 1 : choose_tile(g.M, g.N, g.batch);
    if (code < 0 || code >= kNumTiles || g.M % kTiles[code].bm || g.N % kTiles[code].bn) return hipErrorInvalidValue;
    if (kTiles[code].deep && g.K % 64) return hipErrorInvalidValue;
    if (g.tout == T_F32) return launch_x3<float>(code, p, g.batch, s);
    if (g.tout == T_F64) return launch_x3<double>(code, p, g.batch, s);
    return hipErrorInvalidValue;
}

bool gemm_tile_ok(int code, int tin, int M, int N, int K)
{
    if (code < 1 || code > kNumTiles) return false;
    const TileCfg& t = kTiles[code - 1];
    if (M % t.bm || N % t.bn || (t.x3only && tin != T_BF16X3)) return false;
    if (tin == T_BF16X3) return K % (t.deep ? 64 : 32) == 0;
    const int bk = gemm_k_step(tin, K);
    return K % (t.deep ? 2 * bk : bk) == 0;
}
int gemm_tile_bm(int code) { return (code >= 1 && code <= kNumTiles) ? kTiles[code - 1].bm : 0; }
int gemm_tile_bn(int code) { return (code >= 1 && code <= kNumTiles) ? kTiles[code - 1].bn : 0; }
int gemm_default_tile(int M, int N, int batch) { return choose_tile(M, N, batch) + 1; }
int gemm_k_step(int tin, int K)
{
    if (tin == T_F64) return 16;
    if (tin == T_F32) return K % 32 == 0 ? 32 : 16;
    return K % 64 == 0 ? 64 : 32;
}

hipError_t gemm_grouped(int tin, int tout, bool akm, bool bkm, int tile, int bk, const GemmProb* probs,
                        const int* off, int nprob, int total_tiles, hipStream_t s)
{
    if (nprob <= 0 || total_tiles <= 0) return hipSuccess;
    if (tile < 1 || tile > kNumTiles || kTiles[tile - 1].x3only) return hipErrorInvalidValue;
    const int grid = total_tiles < 2048 ? total_tiles : 2048;
    if (akm && bkm) return grouped_layout<true, true>(tin, tout, tile - 1, bk, probs, off, nprob, grid, s);
    if (akm && !bkm) return grouped_layout<true, false>(tin, tout, tile - 1, bk, probs, off, nprob, grid, s);
    if (!akm && !bkm) return grouped_layout<false, false>(tin, tout, tile - 1, bk, probs, off, nprob, grid, s);
    return grouped_layout<false, true>(tin, tout, tile - 1, bk, probs, off, nprob, grid, s);
}

// Pick the largest tile that divides the problem and still gives >= ~1.5
// workgroups per CU (256 CUs); fall back to the tile with the most blocks.
static int choose_tile(int M, int N, int batch)
{
    int best = -1;
    long best_blocks = -1;
    for (int c = 0; c < kNumTiles; c++) {
        if (kTiles[c].deep || kTiles[c].x3only || M % kTiles[c].bm || N % kTiles[c].bn) continue;
        long blocks = (long)(M / kTiles[c].bm) * (N / kTiles[c].bn) * batch;
        if (blocks >= 384) return c;
        if (blocks > best_blocks) { best_blocks = blocks; best = c; }
    }
    return best;
}

template <typename TI, typename TO, bool AKM, bool BKM>
static hipError_t gemm_typed(const GemmDesc& g, hipStream_t s)
{
    GemmParams p;
    p.A = g.A; p.B = g.B; p.C = g.C;
    p.M = g.M; p.N = g.N; p.K = g.K;
    p.lda = g.lda; p.ldb = g.ldb; p.ldc = g.ldc;
    p.sA = g.sA; p.sB = g.sB; p.sC = g.sC;
    p.alpha = g.alpha; p.accumulate = g.accumulate;
    int code = g.tile > 0 ? g.tile - 1 : choose_tile(g.M, g.N, g.batch);
    if (code < 0 || code >= kNumTiles || kTiles[code].x3only || g.M % kTiles[code].bm || g.N % kTiles[code].bn)
        return hipErrorInvalidValue;
    if (kTiles[code].deep && g.K % (2 * gemm_k_step(g.tin, g.K))) return hipErrorInvalidValue;
    if constexpr (std::is_same<TI, double>::value) {
        if (g.K % 16) return hipErrorInvalidValue;
        return launch_tile<TI, TO, 16, AKM, BKM>(code, p, g.batch, s);
    } else if constexpr (std::is_same<TI, float>::value) {
        if (g.K % 32 == 0) return launch_tile<TI, TO, 32, AKM, BKM>(code, p, g.batch, s);
        if (g.K % 16) return hipErrorInvalidValue;
        return launch_tile<TI, TO, 16, AKM, BKM>(code, p, g.batch, s);
    } else {
        if (g.K % 64 == 0) return launch_tile<TI, TO, 64, AKM, BKM>(code, p, g.batch, s);
        if (g.K % 32) return hipErrorInvalidValue;
        return launch_tile<TI, TO, 32, AKM, BKM>(code, p, g.batch, s);
    }
}

template <bool AKM, bool BKM>
static hipError_t gemm_layout(const GemmDesc& g, hipStream_t s)
{
    if (g.tin == T_F64) {
        if (g.tout == T_F64) return gemm_typed<double, double, AKM, BKM>(g, s);
        return hipErrorInvalidValue;
    }
    if (g.tin == T_F32) {
        if (g.tout == T_F32) return gemm_typed<float, float, AKM, BKM>(g, s);
        if (g.tout == T_F64) return gemm_typed<float, double, AKM, BKM>(g, s);
        return hipErrorInvalidValue;
    }
    if (g.tin == T_BF16) {
        if (g.tout == T_F32) return gemm_typed<bf16, float, AKM, BKM>(g, s);
        if (g.tout == T_F64) return gemm_typed<bf16, double, AKM, BKM>(g, s);
        if (g.tout == T_BF16) return gemm_typed<bf16, bf16, AKM, BKM>(g, s);
    }
    return hipErrorInvalidValue;
}

hipError_t gemm(const GemmDesc& g, hipStream_t s)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.batch <= 0) return hipErrorInvalidValue;
    if (g.M % 32 || g.N % 32 || g.K % 16) return hipErrorInvalidValue;
    if (g.tin == T_BF16X3) return gemm_x3(g, s);
    if (g.a_kmajor && g.b_kmajor) return gemm_layout<true, true>(g, s);
    if (g.a_kmajor && !g.b_kmajor) return gemm_layout<true, false>(g, s);
    if (!g.a_kmajor && !g.b_kmajor) return gemm_layout<false, false>(g, s);
    return gemm_layout<false, true>(g, s);
}

// split-K reduction: C[i*ldc + j] (+)= sum_z W[z*M*N + i*N + j]
template <typename TA, typename TO>
__global__ void __launch_bounds__(256) splitk_reduce_k(const TA* __restrict__ W, int ks, int M, int N,
                                                       TO* __restrict__ C, long ldc, int accumulate)
{
    const long MN = (long)M * N;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < MN; e += (long)gridDim.x * 256) {
        const long i = e / N, j = e - i * N;
        if constexpr (std::is_same<TA, double>::value) {
            double v = 0.0;
            for (int z = 0; z < ks; z++) v += W[z * MN + e];
            TO* cp = C + i * ldc + j;
            if (accumulate) v += to_d(*cp);
            *cp = cvt_from_d<TO>(v);
        } else {
            float v = 0.0f;
            for (int z = 0; z < ks; z++) v += W[z * MN + e];
            TO* cp = C + i * ldc + j;
            if constexpr (std::is_same<TO, double>::value) {
                double dv = (double)v;
                if (accumulate) dv += *cp;
                *cp = dv;
            } else {
                if (accumulate) v += to_f(*cp);
                *cp = cvt_from_f<TO>(v);
            }
        }
    }
}

hipError_t splitk_reduce(int tacc, const void* W, int ks, int M, int N, int tout, void* C, long ldc, int accumulate,
                         hipStream_t s)
{
    long blocks = ((long)M * N + 255) / 256;
    if (blocks > 4096) blocks = 4096;
#define SK_CASE(TAc, TOc, TAt, TOt)                                                                  \
    if (tacc == TAc && tout == TOc) {                                                                \
        splitk_reduce_k<TAt, TOt><<<blocks, 256, 0, s>>>((const TAt*)W, ks, M, N, (TOt*)C, ldc, accumulate); \
        return hipGetLastError();                                                                    \
    }
    SK_CASE(T_F64, T_F64, double, double)
    SK_CASE(T_F32, T_F32, float, float)
    SK_CASE(T_F32, T_F64, float, double)
    SK_CASE(T_F32, T_BF16, float, bf16)
#undef SK_CASE
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// MPO mixing
template <typename TI, typename TO>
__global__ void __launch_bounds__(256) wmix_k(const TI* __restrict__ in, long isx,
                                              TO* __restrict__ out, long osx, int Bn, int nout,
                                              const int* __restrict__ optr,
                                              const long* __restrict__ ooff,
                                              const long* __restrict__ eoff,
                                              const double* __restrict__ eval,
                                              const float* __restrict__ evalf)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= Bn) return;
    const long x = blockIdx.y;
    const TI* ip = in + x * isx + b;
    TO* op = out + x * osx + b;
    for (int o = 0; o < nout; o++) {
        const int e0 = optr[o], e1 = optr[o + 1];
        if constexpr (std::is_same<TI, double>::value) {
            double acc = 0.0;
            for (int e = e0; e < e1; e++) acc += eval[e] * to_d(ip[eoff[e]]);
            op[ooff[o]] = cvt_from_d<TO>(acc);
        } else {
            float acc = 0.0f;
            for (int e = e0; e < e1; e++) acc += evalf[e] * to_f(ip[eoff[e]]);
            op[ooff[o]] = cvt_from_f<TO>(acc);
        }
    }
}

hipError_t wmix(int tin, int tout, const void* in, long isx, void* out, long osx, int X, int Bn,
                const MixTable& t, hipStream_t s)
{
    dim3 grid((Bn + 255) / 256, X);
#define WM_CASE(TIc, TOc, TIt, TOt)                                                         \
    if (tin == TIc && tout == TOc) {                                                        \
        wmix_k<TIt, TOt><<<grid, 256, 0, s>>>((const TIt*)in, isx, (TOt*)out, osx, Bn,      \
                                               t.nout, t.optr, t.ooff, t.eoff, t.eval,      \
                                               t.evalf);                                    \
        return hipGetLastError();                                                           \
    }
    WM_CASE(T_F64, T_F64, double, double)
    WM_CASE(T_F32, T_F32, float, float)
    WM_CASE(T_F32, T_BF16, float, bf16)
    WM_CASE(T_F32, T_F64, float, double)
#undef WM_CASE
    return hipErrorInvalidValue;
}

template <typename TI, typename TX, typename TO, bool X3 = false>
__global__ void __launch_bounds__(256) wmix_idt_k(const TI* __restrict__ in, long isx, const TX* __restrict__ in2,
                                                  long isx2, TO* __restrict__ out, long ldo, long ops,
                                                  double* __restrict__ out2, long osx2, int Bn, int nout,
                                                  const int* __restrict__ optr, const long* __restrict__ ooff,
                                                  const long* __restrict__ eoff, const int* __restrict__ esel,
                                                  const int* __restrict__ osel, const double* __restrict__ eval,
                                                  const float* __restrict__ evalf)
{
    typedef typename std::conditional<std::is_same<TI, double>::value, double, float>::type acc_t;
    __shared__ acc_t tile[64][33];
    const int tb = threadIdx.x & 63, tr = threadIdx.x >> 6;  // b within the tile, row group of 8
    const int b = blockIdx.x * 64 + tb;
    const long x0 = (long)blockIdx.y * 32 + tr * 8;
    const bool bok = b < Bn;
    // writer role: row wb of the tile (b), columns wc..wc+7 (x)
    const int wb = threadIdx.x >> 2, wc = (threadIdx.x & 3) * 8;
    const bool wok = blockIdx.x * 64 + wb < Bn;
    for (int o = 0; o < nout; o++) {
        const int e0 = optr[o], e1 = optr[o + 1];
        acc_t acc[8];
#pragma unroll
        for (int r = 0; r < 8; r++) acc[r] = 0;
        if (bok) {
            for (int e = e0; e < e1; e++) {
                const acc_t v = std::is_same<TI, double>::value ? (acc_t)eval[e] : (acc_t)evalf[e];
                const long off = eoff[e] + b;
                if (esel[e]) {
#pragma unroll
                    for (int r = 0; r < 8; r++) acc[r] += v * to_acc<acc_t>(in2[(x0 + r) * isx2 + off]);
                } else {
#pragma unroll
                    for (int r = 0; r < 8; r++) acc[r] += v * to_acc<acc_t>(in[(x0 + r) * isx + off]);
                }
            }
        }
        if (osel[o]) {  // identity channel of L: straight into y[x][m][b]
            if (bok) {
#pragma unroll
                for (int r = 0; r < 8; r++) out2[(x0 + r) * osx2 + ooff[o] + b] = (double)acc[r];
            }
            continue;
        }
        __syncthreads();  // the previous output's tile has been written out
#pragma unroll
        for (int r = 0; r < 8; r++) tile[tb][tr * 8 + r] = acc[r];
        __syncthreads();
        if (wok) {
            TO* dst = out + ooff[o] + (long)(blockIdx.x * 64 + wb) * ldo + (long)blockIdx.y * 32 + wc;
            if constexpr (X3) {
                // three bfloat16 planes, one 16-byte store per plane (8 values;
                // every offset here is a multiple of 8 elements)
                uint32_t h[3][8];
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    bf16 h0, h1, h2;
                    split3_val((float)tile[wb][wc + c], h0, h1, h2);
                    h[0][c] = h0;
                    h[1][c] = h1;
                    h[2][c] = h2;
                }
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    u32x4 w;
                    w.x = h[q][0] | (h[q][1] << 16);
                    w.y = h[q][2] | (h[q][3] << 16);
                    w.z = h[q][4] | (h[q][5] << 16);
                    w.w = h[q][6] | (h[q][7] << 16);
                    *(u32x4*)(dst + q * ops) = w;
                }
            } else {
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    if constexpr (std::is_same<TO, double>::value) dst[c] = (double)tile[wb][wc + c];
                    else dst[c] = cvt_from_f<TO>((float)tile[wb][wc + c]);
                }
            }
        }
    }
}

hipError_t wmix_idt(int tin, int tx, int tout, const void* in, long isx, const void* in2, long isx2, void* out,
                    long ldo, double* out2, long osx2, int X, int Bn, const MixTable& t, hipStream_t s, long ops)
{
    if (!t.esel || !t.osel || X % 32) return hipErrorInvalidValue;
    dim3 grid((Bn + 63) / 64, X / 32);
    if (tout == T_BF16X3) {
        if (tin != T_F32 || tx != T_F64 || ops <= 0 || ops % 8 || ldo % 8) return hipErrorInvalidValue;
        wmix_idt_k<float, double, bf16, true><<<grid, 256, 0, s>>>((const float*)in, isx, (const double*)in2, isx2,
                                                                 (bf16*)out, ldo, ops, out2, osx2, Bn, t.nout, t.optr,
                                                                 t.ooff, t.eoff, t.esel, t.osel, t.eval, t.evalf);
        return hipGetLastError();
    }
#define WT_CASE(TIc, TXc, TOc, TIt, TXt, TOt)                                                              \
    if (tin == TIc && tx == TXc && tout == TOc) {                                                          \
        wmix_idt_k<TIt, TXt, TOt><<<grid, 256, 0, s>>>((const TIt*)in, isx, (const TXt*)in2, isx2,         \
                                                       (TOt*)out, ldo, 0, out2, osx2, Bn, t.nout, t.optr,  \
                                                       t.ooff, t.eoff, t.esel, t.osel, t.eval, t.evalf);   \
        return hipGetLastError();                                                                          \
    }
    WT_CASE(T_F64, T_F64, T_F64, double, double, double)
    WT_CASE(T_F32, T_F32, T_F32, float, float, float)
    WT_CASE(T_F32, T_BF16, T_BF16, float, bf16, bf16)
#undef WT_CASE
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Stage 1 of the split-bfloat16 Heff apply with the MPO stage fused into its
// epilogue.  A workgroup owns 32 a' x 32 b: its stage-1 tile is rows
// (a', n) for the 32 a' (32 P rows of x) and columns (w', b) for every kept
// right channel w' and the 32 b (the R rows w' D2p + b0 + 0..31), so every
// (n, w') term of the MPO stage for its outputs is in the tile.  The K loop
// is the four-wave loop of gemm_tile_x3 (swizzled LDS, register prefetch);
// then the FP32 tile and the FP64 x slice of the right identity channel go
// to LDS and the mixing sums are formed there: outputs of the kept left
// channels with lanes along a' (32 consecutive a' of one (m, b, w) are one
// 64-byte run of each plane), the left identity channel's outputs with
// lanes along b (coalesced FP64 rows of y).
struct X3MixParams {
    const bf16* X;
    long psX;
    const bf16* R;
    long psR;
    const double* x;
    bf16* T2;
    long ps2, ldo;
    double* y;
    int D1p, D2p;
    const int* optr;
    const int* ocode;
    const int* ecode;
    const float* val;
    int nout, nnz;
};

template <int P, int NR>
__global__ void __launch_bounds__(256) x3_stage1_mix_k(X3MixParams q)
{
    typedef MT<bf16> T;
    typedef T::frag_t frag_t;
    typedef T::acc_t acc_t;
    constexpr int BM = 32 * P, BN = 32 * NR, BK = 32;
    constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
    constexpr int SA = BM * BK, SB = BN * BK;
    constexpr int RA = (BM * 4 + 255) / 256, RB = (BN * 4 + 255) / 256;
    constexpr int TS = BN + 1;  // row stride of the FP32 tile (floats)
    constexpr int XS = 33;      // row stride of the x slice
    constexpr int OPB = 3 * (SA + SB) * 2;
    constexpr int EPB = (BM * TS + BM * XS) * 4;
    constexpr int LB = OPB > EPB ? OPB : EPB;
    __shared__ __attribute__((aligned(16))) unsigned char smem[LB];
    __shared__ int s_optr[33], s_ocode[32], s_ecode[256];
    __shared__ float s_val[256];
    bf16* As = (bf16*)smem;
    bf16* Bs = As + 3 * SA;
    float* Ts = (float*)smem;
    float* Xs = Ts + BM * TS;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int ntm = q.D1p / 32, ntn = q.D2p / 32, nblk = ntm * ntn;
    int bid = blockIdx.x;
    if ((nblk & 7) == 0) bid = (bid & 7) * (nblk >> 3) + (bid >> 3);
    constexpr int GM = 8;
    const int gsz = GM * ntn, g = bid / gsz, first = g * GM;
    const int rows = ntm - first < GM ? ntm - first : GM;
    const int local = bid - g * gsz;
    const int a0 = (first + local % rows) * 32, b0 = (local / rows) * 32;
    const int m0 = a0 * P;
    const long D2p = q.D2p;

    for (int i = tid; i <= q.nout; i += 256) s_optr[i] = q.optr[i];
    for (int i = tid; i < q.nout; i += 256) s_ocode[i] = q.ocode[i];
    for (int i = tid; i < q.nnz; i += 256) {
        s_ecode[i] = q.ecode[i];
        s_val[i] = q.val[i];
    }
    // the x slice of the right identity channel (rows (a', n), 32 b), read
    // now so that its latency hides behind the K loop
    constexpr int XR = BM * 32 / 256;
    float xr[XR];
#pragma unroll
    for (int k = 0; k < XR; k++) {
        const int v = tid + k * 256;
        xr[k] = (float)q.x[(long)(m0 + (v >> 5)) * D2p + b0 + (v & 31)];
    }

    const int wm0 = (wid >> 1) * WM, wn0 = (wid & 1) * WN;
    u32x4 ra[3][RA], rb[3][RB];
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[i][j][r] = 0;
    // operand slices: A rows m0 + t, B rows (t / 32) D2p + b0 + t % 32
    auto gload = [&](int k0) {
#pragma unroll
        for (int pl = 0; pl < 3; pl++) {
#pragma unroll
            for (int r = 0; r < RA; r++) {
                const int v = tid + r * 256;
                if ((BM * 4) % 256 == 0 || v < BM * 4)
                    ra[pl][r] = *(const u32x4*)(q.X + pl * q.psX + (long)(m0 + v / 4) * D2p + k0 + (v % 4) * 8);
            }
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const int v = tid + r * 256;
                if ((BN * 4) % 256 == 0 || v < BN * 4) {
                    const int t = v / 4;
                    const long row = (long)(t >> 5) * D2p + b0 + (t & 31);
                    rb[pl][r] = *(const u32x4*)(q.R + pl * q.psR + row * D2p + k0 + (v % 4) * 8);
                }
            }
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int pl = 0; pl < 3; pl++) {
            x3_sstore<BM, BK>(ra[pl], As + pl * SA, tid);
            x3_sstore<BN, BK>(rb[pl], Bs + pl * SB, tid);
        }
    };
    const int nk = (int)(D2p / BK);
    gload(0);
    sstore();
    __syncthreads();
    for (int t = 0; t < nk; t++) {
        if (t + 1 < nk) gload((t + 1) * BK);
        __builtin_amdgcn_s_setprio(1);
        frag_t a[3][FM], b[3][FN];
#pragma unroll
        for (int pl = 0; pl < 3; pl++) {
#pragma unroll
            for (int i = 0; i < FM; i++)
                a[pl][i] = *(const frag_t*)(As + pl * SA + x3_swz(wm0 + i * 16 + (lane & 15), lane >> 4));
#pragma unroll
            for (int j = 0; j < FN; j++)
                b[pl][j] = *(const frag_t*)(Bs + pl * SB + x3_swz(wn0 + j * 16 + (lane & 15), lane >> 4));
        }
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) {
                acc[i][j] = mma(a[2][i], b[0][j], acc[i][j]);
                acc[i][j] = mma(a[0][i], b[2][j], acc[i][j]);
                acc[i][j] = mma(a[1][i], b[1][j], acc[i][j]);
                acc[i][j] = mma(a[1][i], b[0][j], acc[i][j]);
                acc[i][j] = mma(a[0][i], b[1][j], acc[i][j]);
                acc[i][j] = mma(a[0][i], b[0][j], acc[i][j]);
            }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (t + 1 < nk) {
            sstore();
            __syncthreads();
        }
    }
    // the operand tiles are consumed (barrier above): the LDS takes the FP32
    // stage-1 tile and the x slice (rows (a', n), 32 b)
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                Ts[(wm0 + i * 16 + T::crow(lane, r)) * TS + wn0 + j * 16 + (lane & 15)] = acc[i][j][r];
#pragma unroll
    for (int k = 0; k < XR; k++) {
        const int v = tid + k * 256;
        Xs[(v >> 5) * XS + (v & 31)] = xr[k];
    }
    __syncthreads();
    // sums of the table entries of output o for NV values at rows a_i = al0 +
    // i * astep (a' within the tile) and column bl: an entry's code and value
    // are read once, then its NV source reads are independent
    auto mixsums = [&](int o, int al0, int astep, int bl, float* sum, auto nv) {
        constexpr int NV = decltype(nv)::value;
#pragma unroll
        for (int i = 0; i < NV; i++) sum[i] = 0.f;
        const int e1 = s_optr[o + 1];
        for (int e = s_optr[o]; e < e1; e++) {
            const int ec = s_ecode[e], n = ec & 255;
            const float v = s_val[e];
            const bool fromx = ec >> 16;
            const float* src = fromx ? Xs + n * XS + bl : Ts + n * TS + ((ec >> 8) & 255) * 32 + bl;
            const int rs = (fromx ? XS : TS) * P;  // row step per a'
#pragma unroll
            for (int i = 0; i < NV; i++) sum[i] += v * src[(al0 + i * astep) * rs];
        }
    };
    // kept left channels: T2t[(m, b)][(w, a')] as three planes; a lane forms
    // 8 consecutive a' of one (m, b, w) and stores them as one 16-byte write
    // per plane (4 lanes cover the tile's 32 a'), the two halves of the
    // workgroup take alternate outputs
    {
        const int a8 = (tid & 3) * 8, bl = (tid >> 2) & 31, oh = tid >> 7;
        for (int o = oh; o < q.nout; o += 2) {
            const int oc = s_ocode[o];
            if (oc >> 16) continue;
            float v8[8];
            mixsums(o, a8, 1, bl, v8, std::integral_constant<int, 8>());
            uint32_t h[3][4];
#pragma unroll
            for (int i = 0; i < 8; i += 2) {
                bf16 p0, p1, p2, r0, r1, r2;
                split3_val(v8[i], p0, p1, p2);
                split3_val(v8[i + 1], r0, r1, r2);
                h[0][i / 2] = (uint32_t)p0 | ((uint32_t)r0 << 16);
                h[1][i / 2] = (uint32_t)p1 | ((uint32_t)r1 << 16);
                h[2][i / 2] = (uint32_t)p2 | ((uint32_t)r2 << 16);
            }
            bf16* dst = q.T2 + (long)(oc & 255) * D2p * q.ldo + (long)((oc >> 8) & 255) * q.D1p +
                        (long)(b0 + bl) * q.ldo + a0 + a8;
#pragma unroll
            for (int pl = 0; pl < 3; pl++) {
                u32x4 w;
                w.x = h[pl][0];
                w.y = h[pl][1];
                w.z = h[pl][2];
                w.w = h[pl][3];
                *(u32x4*)(dst + pl * q.ps2) = w;
            }
        }
    }
    // the left identity channel: y[a'][m][b], lanes along b
    {
        const int bl = tid & 31, ag = tid >> 5;
        for (int o = 0; o < q.nout; o++) {
            const int oc = s_ocode[o];
            if (!(oc >> 16)) continue;
            float v4[4];
            mixsums(o, ag, 8, bl, v4, std::integral_constant<int, 4>());
#pragma unroll
            for (int k = 0; k < 4; k++)
                q.y[((long)(a0 + ag + 8 * k) * P + (oc & 255)) * D2p + b0 + bl] = (double)v4[k];
        }
    }
}

bool x3_stage1_mix_ok(int p, int nR, const MixTable& t, int D1p, int D2p)
{
    return t.ocode && t.ecode && t.nout <= 32 && t.nnz <= 256 && D1p > 0 && D2p > 0 && D1p % 32 == 0 &&
           D2p % 32 == 0 && (p == 2 || p == 3) && nR >= 2 && nR <= 5;
}

hipError_t x3_stage1_mix(int p, int nR, const bf16* X, long psX, const bf16* R, long psR, const double* x,
                         bf16* T2, long ps2, long ldo, double* y, int D1p, int D2p, const MixTable& t,
                         hipStream_t s)
{
    if (!x3_stage1_mix_ok(p, nR, t, D1p, D2p)) return hipErrorInvalidValue;
    X3MixParams q;
    q.X = X; q.psX = psX; q.R = R; q.psR = psR; q.x = x; q.T2 = T2; q.ps2 = ps2; q.ldo = ldo; q.y = y;
    q.D1p = D1p; q.D2p = D2p;
    q.optr = t.optr; q.ocode = t.ocode; q.ecode = t.ecode; q.val = t.evalf; q.nout = t.nout; q.nnz = t.nnz;
    const dim3 grid((D1p / 32) * (D2p / 32));
#define X3M_CASE(PP, NN)                                                  \
    if (p == PP && nR == NN) {                                            \
        x3_stage1_mix_k<PP, NN><<<grid, 256, 0, s>>>(q);                  \
        return hipGetLastError();                                         \
    }
    X3M_CASE(2, 2) X3M_CASE(2, 3) X3M_CASE(2, 4) X3M_CASE(2, 5)
    X3M_CASE(3, 2) X3M_CASE(3, 3) X3M_CASE(3, 4) X3M_CASE(3, 5)
#undef X3M_CASE
    return hipErrorInvalidValue;
}

__global__ void __launch_bounds__(256) publish_k(const double* __restrict__ src, int n, double* host_dst,
                                                 unsigned long long* host_flag, unsigned long long seq)
{
    for (int t = threadIdx.x; t < n; t += 256) host_dst[t] = src[t];
    // every wave orders its own stores at system scope before the barrier;
    // the barrier alone is a workgroup-scope fence, so without this the host
    // could see the flag before the values of waves 1-3
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence_system();
        *(volatile unsigned long long*)host_flag = seq;
    }
}

hipError_t publish(const double* src, int n, double* host_dst, unsigned long long* host_flag,
                   unsigned long long seq, hipStream_t s)
{
    if (n < 0 || n > TN_PUBLISH_MAX) return hipErrorInvalidValue;
    publish_k<<<1, 256, 0, s>>>(src, n, host_dst, host_flag, seq);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// elementwise
template <typename TI, typename TO>
__global__ void convert_k(const TI* __restrict__ in, TO* __restrict__ out, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        out[i] = cvt_from_d<TO>(to_d(in[i]));
}

hipError_t convert(int tin, const void* in, int tout, void* out, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
#define CV_CASE(TIc, TOc, TIt, TOt)                                                           \
    if (tin == TIc && tout == TOc) {                                                          \
        convert_k<TIt, TOt><<<blocks, 256, 0, s>>>((const TIt*)in, (TOt*)out, n);             \
        return hipGetLastError();                                                             \
    }
    CV_CASE(T_F64, T_F32, double, float)
    CV_CASE(T_F64, T_BF16, double, bf16)
    CV_CASE(T_F32, T_F64, float, double)
    CV_CASE(T_BF16, T_F64, bf16, double)
    CV_CASE(T_F64, T_F64, double, double)
#undef CV_CASE
    return hipErrorInvalidValue;
}

template <typename TI>
__global__ void split3_k(const TI* __restrict__ in, bf16* __restrict__ out, long n, long ps)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        bf16 h0, h1, h2;
        split3_val(in[i], h0, h1, h2);
        out[i] = h0;
        out[ps + i] = h1;
        out[2 * ps + i] = h2;
    }
}

hipError_t split3(int tin, const void* in, void* out, long n, long ps, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (ps < n) return hipErrorInvalidValue;
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (tin == T_F64) split3_k<double><<<blocks, 256, 0, s>>>((const double*)in, (bf16*)out, n, ps);
    else if (tin == T_F32) split3_k<float><<<blocks, 256, 0, s>>>((const float*)in, (bf16*)out, n, ps);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// x <- a x in place, and the three bfloat16 planes of the scaled values
// (the Lanczos scaling of the next basis vector fused into the apply's first
// pass over it; a x rounds exactly as the separate scaling would)
__global__ void split3_scale_k(double* __restrict__ x, double a, bf16* __restrict__ out, long n, long ps)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double v = a * x[i];
        x[i] = v;
        bf16 h0, h1, h2;
        split3_val(v, h0, h1, h2);
        out[i] = h0;
        out[ps + i] = h1;
        out[2 * ps + i] = h2;
    }
}
hipError_t split3_scale(double* x, double a, void* out, long n, long ps, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (ps < n) return hipErrorInvalidValue;
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    split3_scale_k<<<blocks, 256, 0, s>>>(x, a, (bf16*)out, n, ps);
    return hipGetLastError();
}
// x <- a x in place, and the scaled values converted to TO
template <typename TO>
__global__ void convert_scale_k(double* __restrict__ x, double a, TO* __restrict__ out, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double v = a * x[i];
        x[i] = v;
        out[i] = cvt_from_d<TO>(v);
    }
}
hipError_t convert_scale(double* x, double a, int tout, void* out, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (tout == T_F32) convert_scale_k<float><<<blocks, 256, 0, s>>>(x, a, (float*)out, n);
    else if (tout == T_BF16) convert_scale_k<bf16><<<blocks, 256, 0, s>>>(x, a, (bf16*)out, n);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

__global__ void fill_zero_k(uint32_t* p, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = 0u;
}
hipError_t fill_zero(void* p, long bytes, hipStream_t s)
{
    if (bytes <= 0) return hipSuccess;
    long n = bytes / 4;
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    fill_zero_k<<<blocks, 256, 0, s>>>((uint32_t*)p, n);
    return hipGetLastError();
}
__global__ void fill_value_k(double* p, double v, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = v;
}
hipError_t fill_value(double* p, double v, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    fill_value_k<<<blocks, 256, 0, s>>>(p, v, n);
    return hipGetLastError();
}

// site tensor layouts: reference psi[m + d*(a + D1*b)] <-> internal [D1p][d][D2p]
// (zero padding written on the internal side; grid-stride over the written side)
__global__ void psi_r2i_k(const double* __restrict__ ref, double* __restrict__ out, int d, int D1, int D2, int D1p,
                          int D2p)
{
    const long n = (long)D1p * d * D2p;
    for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int b = (int)(t % D2p);
        const long r = t / D2p;
        const int m = (int)(r % d), a = (int)(r / d);
        out[t] = (a < D1 && b < D2) ? ref[m + (long)d * (a + (long)D1 * b)] : 0.0;
    }
}
__global__ void psi_i2r_k(const double* __restrict__ in, double* __restrict__ ref, int d, int D1, int D2, int D2p)
{
    const long n = (long)d * D1 * D2;
    for (long t = blockIdx.x * 256L + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int m = (int)(t % d);
        const long r = t / d;
        const int a = (int)(r % D1), b = (int)(r / D1);
        ref[t] = in[((long)a * d + m) * D2p + b];
    }
}
hipError_t psi_ref_to_int(const double* ref, double* out, int d, int D1, int D2, int D1p, int D2p, hipStream_t s)
{
    if (d <= 0 || D1 <= 0 || D2 <= 0 || D1p < D1 || D2p < D2) return hipErrorInvalidValue;
    const long n = (long)D1p * d * D2p;
    long blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    psi_r2i_k<<<blocks, 256, 0, s>>>(ref, out, d, D1, D2, D1p, D2p);
    return hipGetLastError();
}
hipError_t psi_int_to_ref(const double* in, double* ref, int d, int D1, int D2, int D2p, hipStream_t s)
{
    if (d <= 0 || D1 <= 0 || D2 <= 0 || D2p < D2) return hipErrorInvalidValue;
    const long n = (long)d * D1 * D2;
    long blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    psi_i2r_k<<<blocks, 256, 0, s>>>(in, ref, d, D1, D2, D2p);
    return hipGetLastError();
}

__global__ void copy3_k(const double* __restrict__ in, long si, long sj, long sk,
                        double* __restrict__ out, int n1, int n2, int n2o)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n2o) return;
    const int j = blockIdx.y, i = blockIdx.z;
    out[((long)i * n1 + j) * n2o + k] = (k < n2) ? in[i * si + j * sj + k * sk] : 0.0;
}
hipError_t copy3(const double* in, long si, long sj, long sk, double* out, int n0, int n1,
                 int n2, int n2o, hipStream_t s)
{
    if (n0 <= 0 || n1 <= 0 || n2o <= 0) return hipSuccess;
    dim3 grid((n2o + 255) / 256, n1, n0);
    copy3_k<<<grid, 256, 0, s>>>(in, si, sj, sk, out, n1, n2, n2o);
    return hipGetLastError();
}

__global__ void copy2_k(const double* __restrict__ in, long si0, long si1, int r, int c,
                        double* __restrict__ out, long so0, long so1, int co)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= co) return;
    const int i = blockIdx.y;
    out[i * so0 + j * so1] = (i < r && j < c) ? in[i * si0 + j * si1] : 0.0;
}
hipError_t copy2(const double* in, long si0, long si1, int r, int c, double* out, long so0,
                 long so1, int ro, int co, hipStream_t s)
{
    if (ro <= 0 || co <= 0) return hipSuccess;
    dim3 grid((co + 255) / 256, ro);
    copy2_k<<<grid, 256, 0, s>>>(in, si0, si1, r, c, out, so0, so1, co);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Lanczos vector kernels
#define MD_BLOCK 256
#define MD_EPT 8
#define MD_MAXK 512

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

long mdot_part_size(long n, int k)
{
    long blocks = (n + MD_BLOCK * MD_EPT - 1) / (MD_BLOCK * MD_EPT);
    return blocks * (long)k;
}

// stage 1: part[i*G + blk] = sum over the block's chunk of V[i]*w
__global__ void __launch_bounds__(MD_BLOCK) mdot1_k(const double* __restrict__ V, long ldv, int k,
                                                    const double* __restrict__ w, long n,
                                                    double* __restrict__ part)
{
    __shared__ double red[MD_MAXK][MD_BLOCK / 64];
    const long base = (long)blockIdx.x * MD_BLOCK * MD_EPT + threadIdx.x;
    double wr[MD_EPT];
#pragma unroll
    for (int e = 0; e < MD_EPT; e++) {
        long idx = base + (long)e * MD_BLOCK;
        wr[e] = idx < n ? w[idx] : 0.0;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = 0; i < k; i++) {
        const double* vi = V + (long)i * ldv;
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < MD_EPT; e++) {
            long idx = base + (long)e * MD_BLOCK;
            if (idx < n) acc += vi[idx] * wr[e];
        }
        acc = wave_sum(acc);
        if (lane == 0) red[i][wid] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += MD_BLOCK) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < MD_BLOCK / 64; q++) s += red[i][q];
        part[(long)i * gridDim.x + blockIdx.x] = s;
    }
}

// stage 2: out[i] = sum_b part[i*G + b]  (one block per i, fixed order)
__global__ void __launch_bounds__(256) mdot2_k(const double* __restrict__ part, int G,
                                               double* __restrict__ out)
{
    __shared__ double red[4];
    const int i = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < G; b += 256) acc += part[(long)i * G + b];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[i] = red[0] + red[1] + red[2] + red[3];
}

hipError_t mdot(const double* V, long ldv, int k, const double* w, long n, double* out,
                double* part, int part_cap, hipStream_t s)
{
    if (k <= 0) return hipSuccess;
    long G = (n + MD_BLOCK * MD_EPT - 1) / (MD_BLOCK * MD_EPT);
    for (int i0 = 0; i0 < k; i0 += MD_MAXK) {
        const int kc = (k - i0) < MD_MAXK ? (k - i0) : MD_MAXK;
        if (G * kc > part_cap) return hipErrorInvalidValue;
        mdot1_k<<<G, MD_BLOCK, 0, s>>>(V + (long)i0 * ldv, ldv, kc, w, n, part);
        mdot2_k<<<kc, 256, 0, s>>>(part, (int)G, out + i0);
    }
    return hipGetLastError();
}

// Fused Lanczos kernels.  A block owns a chunk of MD_BLOCK * MD_EPT entries,
// each thread MD_EPT / 2 pairs (16-byte loads, consecutive lanes on
// consecutive pairs).  update_norm visits the chunks in the reverse order of
// step_dots, so its first chunks of the basis are the ones step_dots read
// last (still in the last-level cache when the basis exceeds it).
#define LZ_PAIRS (MD_EPT / 2)
__global__ void __launch_bounds__(MD_BLOCK) lz_step_dots_k(double* __restrict__ w, const double* __restrict__ V,
                                                           long ldv, int k, const double* __restrict__ alpha,
                                                           double beta, long n, double* __restrict__ part)
{
    __shared__ double red[MD_MAXK][MD_BLOCK / 64];
    const long c0 = (long)blockIdx.x * MD_BLOCK * MD_EPT;
    const double a = alpha[0];
    const double* vj = V + (long)(k - 1) * ldv;
    const double* vm = k >= 2 ? V + (long)(k - 2) * ldv : nullptr;
    double2 wr[LZ_PAIRS];
#pragma unroll
    for (int e = 0; e < LZ_PAIRS; e++) {
        const long idx = c0 + 2 * ((long)e * MD_BLOCK + threadIdx.x);
        if (idx < n) {
            double2 x = *(const double2*)(w + idx);
            const double2 p = *(const double2*)(vj + idx);
            x.x -= a * p.x;
            x.y -= a * p.y;
            if (vm) {
                const double2 q = *(const double2*)(vm + idx);
                x.x -= beta * q.x;
                x.y -= beta * q.y;
            }
            *(double2*)(w + idx) = x;
            wr[e] = x;
        } else {
            wr[e] = make_double2(0.0, 0.0);
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = 0; i < k; i++) {
        const double* vi = V + (long)i * ldv;
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < LZ_PAIRS; e++) {
            const long idx = c0 + 2 * ((long)e * MD_BLOCK + threadIdx.x);
            if (idx < n) {
                const double2 p = *(const double2*)(vi + idx);
                acc += p.x * wr[e].x + p.y * wr[e].y;
            }
        }
        acc = wave_sum(acc);
        if (lane == 0) red[i][wid] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += MD_BLOCK) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < MD_BLOCK / 64; q++) sum += red[i][q];
        part[(long)i * gridDim.x + blockIdx.x] = sum;
    }
}

__global__ void __launch_bounds__(MD_BLOCK) lz_update_norm_k(double* __restrict__ w, const double* __restrict__ V,
                                                             long ldv, int k, const double* __restrict__ coef,
                                                             long n, double* __restrict__ part)
{
    __shared__ double c[MD_MAXK];
    __shared__ double red[MD_BLOCK / 64];
    for (int i = threadIdx.x; i < k; i += MD_BLOCK) c[i] = coef[i];
    __syncthreads();
    const int blk = gridDim.x - 1 - blockIdx.x;
    const long c0 = (long)blk * MD_BLOCK * MD_EPT;
    double2 acc[LZ_PAIRS];
#pragma unroll
    for (int e = 0; e < LZ_PAIRS; e++) {
        const long idx = c0 + 2 * ((long)e * MD_BLOCK + threadIdx.x);
        acc[e] = idx < n ? *(const double2*)(w + idx) : make_double2(0.0, 0.0);
    }
    for (int i = 0; i < k; i++) {
        const double* vi = V + (long)i * ldv;
        const double ci = c[i];
#pragma unroll
        for (int e = 0; e < LZ_PAIRS; e++) {
            const long idx = c0 + 2 * ((long)e * MD_BLOCK + threadIdx.x);
            if (idx < n) {
                const double2 p = *(const double2*)(vi + idx);
                acc[e].x -= ci * p.x;
                acc[e].y -= ci * p.y;
            }
        }
    }
    double sq = 0.0;
#pragma unroll
    for (int e = 0; e < LZ_PAIRS; e++) {
        const long idx = c0 + 2 * ((long)e * MD_BLOCK + threadIdx.x);
        if (idx < n) {
            *(double2*)(w + idx) = acc[e];
            sq += acc[e].x * acc[e].x + acc[e].y * acc[e].y;
        }
    }
    sq = wave_sum(sq);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < MD_BLOCK / 64; q++) sum += red[q];
        part[blk] = sum;
    }
}

hipError_t lz_step_dots(double* w, const double* V, long ldv, int k, const double* alpha, double beta, long n,
                        double* out, double* part, long part_cap, hipStream_t s)
{
    const long G = (n + MD_BLOCK * MD_EPT - 1) / (MD_BLOCK * MD_EPT);
    if (k < 1 || k > MD_MAXK || (n & 1) || G * k > part_cap) return hipErrorInvalidValue;
    lz_step_dots_k<<<G, MD_BLOCK, 0, s>>>(w, V, ldv, k, alpha, beta, n, part);
    mdot2_k<<<k, 256, 0, s>>>(part, (int)G, out);
    return hipGetLastError();
}

hipError_t lz_update_norm(double* w, const double* V, long ldv, int k, const double* coef, long n, double* out,
                          double* part, long part_cap, hipStream_t s)
{
    const long G = (n + MD_BLOCK * MD_EPT - 1) / (MD_BLOCK * MD_EPT);
    if (k < 1 || k > MD_MAXK || (n & 1) || G > part_cap) return hipErrorInvalidValue;
    lz_update_norm_k<<<G, MD_BLOCK, 0, s>>>(w, V, ldv, k, coef, n, part);
    mdot2_k<<<1, 256, 0, s>>>(part, (int)G, out);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) gemv_acc_k(const double* __restrict__ V, long ldv, int k,
                                                  const double* __restrict__ coef,
                                                  double coef_scale, double beta_w,
                                                  double* __restrict__ w, long n)
{
    __shared__ double c[MD_MAXK];
    for (int i = threadIdx.x; i < k; i += 256) c[i] = coef[i] * coef_scale;
    __syncthreads();
    for (long idx = blockIdx.x * 256L + threadIdx.x; idx < n; idx += (long)gridDim.x * 256) {
        double acc = (beta_w == 0.0) ? 0.0 : beta_w * w[idx];
        for (int i = 0; i < k; i++) acc += c[i] * V[(long)i * ldv + idx];
        w[idx] = acc;
    }
}
hipError_t gemv_acc(const double* V, long ldv, int k, const double* coef, double coef_scale,
                    double beta_w, double* w, long n, hipStream_t s)
{
    long blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    if (k <= 0) {
        if (beta_w != 1.0) gemv_acc_k<<<blocks, 256, 0, s>>>(V, ldv, 0, coef, coef_scale, beta_w, w, n);
        return hipGetLastError();
    }
    for (int i0 = 0; i0 < k; i0 += MD_MAXK) {
        const int kc = (k - i0) < MD_MAXK ? (k - i0) : MD_MAXK;
        gemv_acc_k<<<blocks, 256, 0, s>>>(V + (long)i0 * ldv, ldv, kc, coef + i0, coef_scale,
                                          i0 == 0 ? beta_w : 1.0, w, n);
    }
    return hipGetLastError();
}

// max |x| over n entries: stage 1 per block, stage 2 one block; out[0] = max
__global__ void __launch_bounds__(256) absmax1_k(const double* __restrict__ x, long n,
                                                 double* __restrict__ part)
{
    __shared__ double red[4];
    double m = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) m = fmax(m, fabs(x[i]));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}
__global__ void __launch_bounds__(256) absmax2_k(const double* __restrict__ part, int G,
                                                 double* __restrict__ out)
{
    __shared__ double red[4];
    double m = 0.0;
    for (int i = threadIdx.x; i < G; i += 256) m = fmax(m, part[i]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}
hipError_t absmax(const double* x, long n, double* out, double* part, int part_cap, hipStream_t s)
{
    if (n <= 0) return hipErrorInvalidValue;
    long G = (n + 255) / 256;
    if (G > 1024) G = 1024;
    if (G > part_cap) G = part_cap;
    if (G < 1) return hipErrorInvalidValue;
    absmax1_k<<<G, 256, 0, s>>>(x, n, part);
    absmax2_k<<<1, 256, 0, s>>>(part, (int)G, out);
    return hipGetLastError();
}

__global__ void scale_by_inv_k(double* v, const double* den, long n)
{
    const double f = 1.0 / den[0];
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) v[i] *= f;
}
hipError_t scale_by_inv(double* v, const double* den, long n, hipStream_t s)
{
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    scale_by_inv_k<<<blocks, 256, 0, s>>>(v, den, n);
    return hipGetLastError();
}
__global__ void axpby_k(double a, const double* x, double b, double* y, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        y[i] = a * x[i] + b * y[i];
}
hipError_t axpby(double a, const double* x, double b, double* y, long n, hipStream_t s)
{
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    axpby_k<<<blocks, 256, 0, s>>>(a, x, b, y, n);
    return hipGetLastError();
}

// w -= alpha[0] v + beta u   (the three-term Lanczos step; alpha on the device)
__global__ void three_term_k(double* __restrict__ w, const double* __restrict__ v, const double* __restrict__ u,
                             const double* __restrict__ alpha, double beta, long n)
{
    const double a = alpha[0];
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        double r = w[i] - a * v[i];
        if (u) r -= beta * u[i];
        w[i] = r;
    }
}
hipError_t three_term(double* w, const double* v, const double* u, const double* alpha, double beta, long n,
                      hipStream_t s)
{
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    three_term_k<<<blocks, 256, 0, s>>>(w, v, u, alpha, beta, n);
    return hipGetLastError();
}

__global__ void scale_copy_k(double a, const double* x, double* y, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) y[i] = a * x[i];
}
hipError_t scale_copy(double a, const double* x, double* y, long n, hipStream_t s)
{
    long blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    scale_copy_k<<<blocks, 256, 0, s>>>(a, x, y, n);
    return hipGetLastError();
}

__global__ void dcopy_scalar_k(const double* src, double* dst)
{
    if (threadIdx.x == 0) dst[0] = src[0];
}
hipError_t dcopy_scalar(const double* src, double* dst, hipStream_t s)
{
    dcopy_scalar_k<<<1, 64, 0, s>>>(src, dst);
    return hipGetLastError();
}
__global__ void dsqrt_scalar_k(const double* src, double* dst)
{
    if (threadIdx.x == 0) dst[0] = sqrt(fmax(src[0], 0.0));
}
hipError_t dsqrt_scalar(const double* src, double* dst, hipStream_t s)
{
    dsqrt_scalar_k<<<1, 64, 0, s>>>(src, dst);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// one-sided block Jacobi SVD: gridDim.y workgroups per column-block pair (2 x
// 16 columns).  Each forms the pair's 32x32 Gram matrix from the source
// iterate Xs, diagonalises it in LDS by cyclic parallel Jacobi and applies the
// rotation to its own chunk of rows: Xs -> Xd, and V in place.  The Gram
// matrix and the rotation are computed the same way by every workgroup of the
// pair (same data, same order: bit-identical), so the q/32 pairs of a round
// spread over up to 16 times as many CUs; reading Xs and writing Xd keeps a
// workgroup from seeing rows another has already rotated.  A pair that is
// already orthogonal copies its rows to Xd.
#define JB 16
#define JW 32
#define JR 128  // rows per Gram chunk

__global__ void __launch_bounds__(256) jacobi_k(const double* __restrict__ Xs, double* __restrict__ Xd, long ldx,
                                                int p, double* __restrict__ V, long ldv, int qv,
                                                const int* __restrict__ pairs, double tol,
                                                int* __restrict__ flag, int inner)
{
    // even row strides: column pairs are 16-byte aligned for double2 reads
    __shared__ double G[JW][JW + 1];
    __shared__ __align__(16) double J[JW][JW + 2];
    __shared__ __align__(16) double tile[JR][JW + 2];
    __shared__ double rc[JW / 2], rs[JW / 2];
    __shared__ int rp[JW / 2], rq[JW / 2];
    __shared__ int any_rot;
    __shared__ int sh_off;
    const int tid = threadIdx.x;
    const int pb = pairs[2 * blockIdx.x], qb = pairs[2 * blockIdx.x + 1];
    auto colidx = [&](int j) { return j < JB ? pb * JB + j : qb * JB + (j - JB); };

    // 1. Gram matrix over row chunks of JR rows staged in LDS (all loads of a
    // chunk in flight together: few load-then-barrier rounds per launch), on
    // the FP64 matrix cores: wave w forms the 16 x 16 block (w >> 1, w & 1)
    // of G, one v_mfma_f64_16x16x4 per 4 rows (lane l reads row l/16 of the
    // step and column l%16 of each of its two column blocks).
    const int lane = tid & 63, wv = tid >> 6;
    const int wi = wv >> 1, wj = wv & 1;
    v4d gacc = {0.0, 0.0, 0.0, 0.0};
    for (int r0 = 0; r0 < p; r0 += JR) {
        double ld[JR * JW / 256];
#pragma unroll
        for (int u = 0; u < JR * JW / 256; u++) {
            const int idx = tid + 256 * u;
            const int row = idx % JR, c = idx / JR;
            const int gr = r0 + row;
            ld[u] = (gr < p) ? Xs[(long)colidx(c) * ldx + gr] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < JR * JW / 256; u++) {
            const int idx = tid + 256 * u;
            tile[idx % JR][idx / JR] = ld[u];
        }
        __syncthreads();
#pragma unroll 8
        for (int k0 = 0; k0 < JR; k0 += 4) {
            const double a = tile[k0 + (lane >> 4)][16 * wi + (lane & 15)];
            const double b = tile[k0 + (lane >> 4)][16 * wj + (lane & 15)];
            gacc = mma(a, b, gacc);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; r++) G[16 * wi + (lane >> 4) + 4 * r][16 * wj + (lane & 15)] = gacc[r];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int e = tid + 256 * u;
        J[e >> 5][e & 31] = ((e >> 5) == (e & 31)) ? 1.0 : 0.0;
    }
    if (tid == 0) sh_off = 0;
    __syncthreads();
    // convergence measure of this pair (full 32x32 off-diagonal)
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int e = tid + 256 * u;
        const int gi = e >> 5, gj = e & 31;
        if (gi < gj) {
            const double dd = G[gi][gi] * G[gj][gj];
            if (dd > 1e-300 && fabs(G[gi][gj]) > tol * sqrt(dd)) sh_off = 1;
        }
    }
    __syncthreads();
    // this workgroup's rows of the iterate and of V
    const int nch = gridDim.y, ch = blockIdx.y;
    const int xrows = (p + nch - 1) / nch, vrows = (qv + nch - 1) / nch;
    const int xr0 = ch * xrows, xr1 = min(p, xr0 + xrows);
    const int vr0 = ch * vrows, vr1 = min(qv, vr0 + vrows);
    if (sh_off == 0) {  // already orthogonal: nothing to rotate
        for (int e = tid; e < JW * (xr1 > xr0 ? xr1 - xr0 : 0); e += 256) {
            const long off = (long)colidx(e / (xr1 - xr0)) * ldx + xr0 + e % (xr1 - xr0);
            Xd[off] = Xs[off];
        }
        return;
    }
    if (tid == 0) flag[0] = 1;

    // 2. cyclic parallel Jacobi on G (tournament ordering, 31 rounds per sweep)
    for (int sweep = 0; sweep < inner; sweep++) {
        if (tid == 0) any_rot = 0;
        __syncthreads();
        for (int round = 0; round < JW - 1; round++) {
            if (tid < JW / 2) {
                auto elem = [&](int kpos) { return kpos == 0 ? 0 : ((kpos - 1 + round) % (JW - 1)) + 1; };
                int a = elem(tid), b = elem(JW - 1 - tid);
                int P = a < b ? a : b, Q = a < b ? b : a;
                double app = G[P][P], aqq = G[Q][Q], apq = G[P][Q];
                double c = 1.0, s = 0.0;
                if (fabs(apq) > 1e-300 && fabs(apq) > 1e-16 * sqrt(fabs(app * aqq))) {
                    double theta = (aqq - app) / (2.0 * apq);
                    double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    c = 1.0 / sqrt(t * t + 1.0);
                    s = t * c;
                    any_rot = 1;
                }
                rc[tid] = c;
                rs[tid] = s;
                rp[tid] = P;
                rq[tid] = Q;
            }
            __syncthreads();
            // G <- R^T G R in one phase: the round's pairs are disjoint, so G
            // splits into 16 x 16 independent 2 x 2 blocks (rows P1, Q1 of
            // pair bi, columns P2, Q2 of pair bj), one per thread
            {
                const int bi = tid >> 4, bj = tid & 15;
                const int P1 = rp[bi], Q1 = rq[bi], P2 = rp[bj], Q2 = rq[bj];
                const double c1 = rc[bi], s1 = rs[bi], c2 = rc[bj], s2 = rs[bj];
                const double g00 = G[P1][P2], g01 = G[P1][Q2], g10 = G[Q1][P2], g11 = G[Q1][Q2];
                const double h00 = c1 * g00 - s1 * g10, h01 = c1 * g01 - s1 * g11;
                const double h10 = s1 * g00 + c1 * g10, h11 = s1 * g01 + c1 * g11;
                G[P1][P2] = c2 * h00 - s2 * h01;
                G[P1][Q2] = s2 * h00 + c2 * h01;
                G[Q1][P2] = c2 * h10 - s2 * h11;
                G[Q1][Q2] = s2 * h10 + c2 * h11;
            }
            // J <- J R
            for (int e = tid; e < (JW / 2) * JW; e += 256) {
                const int pi = e / JW, row = e % JW;
                const int P = rp[pi], Q = rq[pi];
                const double c = rc[pi], s = rs[pi];
                const double jp = J[row][P], jq = J[row][Q];
                J[row][P] = c * jp - s * jq;
                J[row][Q] = s * jp + c * jq;
            }
            __syncthreads();
        }
        if (any_rot == 0) break;
        __syncthreads();
    }

    // 3. X[:, cols] <- X[:, cols] J ; V[:, cols] <- V[:, cols] J (J read in
    // 16-byte column pairs, the same for every lane)
    for (int row = xr0 + tid; row < xr1; row += 256) {
        double x[JW];
#pragma unroll
        for (int c = 0; c < JW; c++) x[c] = Xs[(long)colidx(c) * ldx + row];
#pragma unroll 2
        for (int c2 = 0; c2 < JW; c2 += 2) {
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int c = 0; c < JW; c++) {
                const double2 jj = *(const double2*)&J[c][c2];
                s0 = fma(x[c], jj.x, s0);
                s1 = fma(x[c], jj.y, s1);
            }
            Xd[(long)colidx(c2) * ldx + row] = s0;
            Xd[(long)colidx(c2 + 1) * ldx + row] = s1;
        }
    }
    for (int row = vr0 + tid; row < vr1; row += 256) {
        double x[JW];
#pragma unroll
        for (int c = 0; c < JW; c++) x[c] = V[(long)colidx(c) * ldv + row];
#pragma unroll 2
        for (int c2 = 0; c2 < JW; c2 += 2) {
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int c = 0; c < JW; c++) {
                const double2 jj = *(const double2*)&J[c][c2];
                s0 = fma(x[c], jj.x, s0);
                s1 = fma(x[c], jj.y, s1);
            }
            V[(long)colidx(c2) * ldv + row] = s0;
            V[(long)colidx(c2 + 1) * ldv + row] = s1;
        }
    }
}

hipError_t jacobi_step(const double* Xs, double* Xd, long ldx, int p, double* V, long ldv, int qv,
                       const int* pairs, int npairs, int nchunk, double tol, int* flag, hipStream_t s, int inner)
{
    if (npairs <= 0) return hipSuccess;
    if (nchunk < 1 || nchunk > 64 || Xs == Xd) return hipErrorInvalidValue;
    jacobi_k<<<dim3(npairs, nchunk), 256, 0, s>>>(Xs, Xd, ldx, p, V, ldv, qv, pairs, tol, flag, inner);
    return hipGetLastError();
}

// Householder QR of one 32-column panel (the column loop of LAPACK dgeqr2
// with the forward, columnwise T of dlarft).  A is column-major (column j at
// A + j*lda) with pp rows; the panel is columns c0..c0+31, rows c0..pp-1.
// Column c = c0 + j has its diagonal in row r = c0 + j: H_j = I - tau v v^T,
// v_r = 1, maps rows r.. of the column to beta e_r, so R(r, c) = beta and the
// rows below become zero (a column whose part below the diagonal is zero, or
// underflows to zero when squared, gets tau = 0).  v goes to column j of V
// (rows c0..pp-1, zero above r) and the upper-triangular T (32 x 32,
// row-major) satisfies H_0 H_1 ... H_31 = I - V T V^T.
//
// One workgroup.  Per column three phases, a barrier after each: (a) all 256
// threads form v and the reflected column; (b) the dot products of v with the
// 32 panel columns (the earlier v's for k < j, for T): 8 groups of 32 lanes,
// group g owns columns 4g..4g+3 and each lane every 32nd row, so a thread
// keeps 4 sums and its loads are all in flight together, and the 32 lanes
// reduce by shuffles; (c) H_j on the columns right of j (same layout; the
// others are stored back unchanged so the loads stay unconditional), the next
// column's squared norm below its diagonal, and column j of T.
__global__ void __launch_bounds__(256) hqr_panel_k(double* __restrict__ A, long lda, int pp, int c0,
                                                    double* __restrict__ V, long ldv, double* __restrict__ T)
{
    __shared__ double fin[34];  // [0, 32): dot products; [32]: norm^2 below the diagonal; [33]: diagonal
    __shared__ double Ts[32][33];
    __shared__ double red[4];
    const int tid = threadIdx.x;
    const int g = tid >> 5, rl = tid & 31;
    for (int e = tid; e < 32 * 33; e += 256) Ts[e / 33][e % 33] = 0.0;
    {
        const double* col = A + (long)c0 * lda;
        double s2 = 0.0;
        for (int i = c0 + 1 + tid; i < pp; i += 256) s2 += col[i] * col[i];
        s2 = wave_sum(s2);
        if ((tid & 63) == 0) red[tid >> 6] = s2;
        if (tid == 0) fin[33] = col[c0];
    }
    __syncthreads();
    if (tid == 0) fin[32] = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    for (int j = 0; j < 32; j++) {
        const int r = c0 + j;
        double* col = A + (long)r * lda;
        const double alpha = fin[33], sig2 = fin[32];
        double tau = 0.0, beta = alpha, scl = 0.0;
        if (sig2 > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + sig2), alpha);
            tau = (beta - alpha) / beta;
            scl = 1.0 / (alpha - beta);
        }
        // (a) v and the reflected column
        double* vj = V + (long)j * ldv;
        for (int i = c0 + tid; i < pp; i += 256) {
            double vi = 0.0;
            if (i >= r) {
                vi = i == r ? 1.0 : col[i] * scl;
                col[i] = i == r ? beta : 0.0;
            }
            vj[i] = vi;
        }
        __syncthreads();
        // (b) d_k = src_k . v over rows >= r (src_k = v_k for k < j, column
        // c0 + k otherwise; d_j is not used)
        const double* src[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int k = 4 * g + e;
            src[e] = k < j ? V + (long)k * ldv : A + (long)(c0 + k) * lda;
        }
        const int i0 = c0 + rl + 32 * ((r - c0) >> 5);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int i = i0; i < pp; i += 32) {
            const double vi = i >= r ? vj[i] : 0.0;
#pragma unroll
            for (int e = 0; e < 4; e++) acc[e] += src[e][i] * vi;
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
#pragma unroll
            for (int o = 16; o >= 1; o >>= 1) acc[e] += __shfl_xor(acc[e], o, 32);
        }
        if (rl == 0) {
#pragma unroll
            for (int e = 0; e < 4; e++) fin[4 * g + e] = acc[e];
        }
        __syncthreads();
        // (c) H_j on the columns right of j; the next column's norm; T
        double f[4];
#pragma unroll
        for (int e = 0; e < 4; e++) f[e] = 4 * g + e > j ? tau * fin[4 * g + e] : 0.0;
        double s2 = 0.0;
        const int kn = j + 1;
        const bool own_next = kn < 32 && g == (kn >> 2);
        // every lane has the same number of rows (pp - c0 is a multiple of
        // 32); rows above r keep their values (v is zero there).  Blocks of
        // four rows: all loads, then the updates, then all stores, since a
        // load cannot move above a store the compiler cannot tell apart.
        const int nrow = ((pp - c0) >> 5) - ((r - c0) >> 5);
        auto rows = [&](int t, int nu) {
            double x[4][4], w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (u >= nu) break;
                const int i = i0 + 32 * (t + u);
                w[u] = i >= r ? vj[i] : 0.0;
#pragma unroll
                for (int e = 0; e < 4; e++) x[u][e] = A[(long)(c0 + 4 * g + e) * lda + i];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (u >= nu) break;
                const int i = i0 + 32 * (t + u);
                double y = 0.0;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    x[u][e] -= f[e] * w[u];
                    A[(long)(c0 + 4 * g + e) * lda + i] = x[u][e];
                    y = 4 * g + e == kn ? x[u][e] : y;
                }
                s2 += i > r + 1 ? y * y : 0.0;
                if (own_next && i == r + 1) fin[33] = y;
            }
        };
        int t = 0;
        for (; t + 4 <= nrow; t += 4) rows(t, 4);
        for (; t < nrow; t++) rows(t, 1);
        if (own_next) {
#pragma unroll
            for (int o = 16; o >= 1; o >>= 1) s2 += __shfl_xor(s2, o, 32);
            if (rl == 0) fin[32] = s2;
        }
        if (tid < j) {
            double t = 0.0;
            for (int l = tid; l < j; l++) t += Ts[tid][l] * fin[l];
            Ts[tid][j] = -tau * t;
        } else if (tid == j) {
            Ts[j][j] = tau;
        }
        __syncthreads();
    }
    for (int e = tid; e < 32 * 32; e += 256) T[e] = Ts[e >> 5][e & 31];
}
hipError_t hqr_panel(double* A, long lda, int pp, int c0, double* V, long ldv, double* T, hipStream_t s)
{
    if (c0 < 0 || c0 + 32 > pp || lda < pp || ldv < pp) return hipErrorInvalidValue;
    hqr_panel_k<<<1, 256, 0, s>>>(A, lda, pp, c0, V, ldv, T);
    return hipGetLastError();
}

// m[i] = max over j != i (j < n) of |G_ij| / sqrt(G_ii G_jj), the largest
// cosine between column i and another column of the matrix whose Gram matrix
// G is (row-major, ld ldg); pairs with G_ii G_jj <= 1e-300 are skipped.
// One workgroup per row.
__global__ void __launch_bounds__(256) max_cos_k(const double* __restrict__ G, long ldg, int n,
                                                  double* __restrict__ m)
{
    __shared__ double red[4];
    const int i = blockIdx.x;
    const double gi = G[(long)i * ldg + i];
    double best = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        if (j == i) continue;
        const double d = gi * G[(long)j * ldg + j];
        if (d > 1e-300) best = fmax(best, fabs(G[(long)i * ldg + j]) / sqrt(d));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) best = fmax(best, __shfl_xor(best, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) m[i] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}
hipError_t max_cos(const double* G, long ldg, int n, double* m, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    max_cos_k<<<n, 256, 0, s>>>(G, ldg, n, m);
    return hipGetLastError();
}

// out[j*ldo + i] = in[perm[j]*ldi + i] (i < rows, j < ncols): a column gather
__global__ void gather_cols_k(const double* __restrict__ in, long ldi, double* __restrict__ out, long ldo,
                              const int* __restrict__ perm, int rows)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i < rows) out[(long)j * ldo + i] = in[(long)perm[j] * ldi + i];
}
hipError_t gather_cols(const double* in, long ldi, double* out, long ldo, const int* perm, int rows, int ncols,
                       hipStream_t s)
{
    if (rows <= 0 || ncols <= 0) return hipSuccess;
    gather_cols_k<<<dim3((rows + 255) / 256, ncols), 256, 0, s>>>(in, ldi, out, ldo, perm, rows);
    return hipGetLastError();
}
// X[j*ldx + perm[i]] = R[i*ldr + j] (i, j < n): R^T with its rows permuted
__global__ void tr_scatter_k(const double* __restrict__ R, long ldr, double* __restrict__ X, long ldx,
                             const int* __restrict__ perm, int n)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (j < n) X[(long)j * ldx + perm[i]] = R[(long)i * ldr + j];
}
hipError_t tr_scatter(const double* R, long ldr, double* X, long ldx, const int* perm, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    tr_scatter_k<<<dim3((n + 255) / 256, n), 256, 0, s>>>(R, ldr, X, ldx, perm, n);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) col_norms_k(const double* __restrict__ X, long ldx, int p,
                                                   double* __restrict__ s)
{
    __shared__ double red[4];
    const double* x = X + (long)blockIdx.x * ldx;
    double acc = 0.0;
    for (int i = threadIdx.x; i < p; i += 256) acc += x[i] * x[i];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) s[blockIdx.x] = sqrt(red[0] + red[1] + red[2] + red[3]);
}
hipError_t col_norms(const double* X, long ldx, int p, int q, double* s, hipStream_t st)
{
    if (q <= 0) return hipSuccess;
    col_norms_k<<<q, 256, 0, st>>>(X, ldx, p, s);
    return hipGetLastError();
}

__global__ void set_int_k(int* p, int v)
{
    if (threadIdx.x == 0) p[0] = v;
}
hipError_t set_int(int* p, int v, hipStream_t s)
{
    set_int_k<<<1, 64, 0, s>>>(p, v);
    return hipGetLastError();
}

__global__ void set_identity_k(double* V, long ldv, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) V[(long)j * ldv + i] = (i == j) ? 1.0 : 0.0;
}
hipError_t set_identity(double* V, long ldv, int n, hipStream_t s)
{
    dim3 grid((n + 255) / 256, n);
    set_identity_k<<<grid, 256, 0, s>>>(V, ldv, n);
    return hipGetLastError();
}

__global__ void svd_gather1_k(const double* __restrict__ X, long ldx, int p,
                              const int* __restrict__ perm, const double* __restrict__ scl,
                              double* __restrict__ out, long s0, long s1)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= p) return;
    out[(long)i * s0 + (long)j * s1] = X[(long)perm[j] * ldx + i] * scl[j];
}
__global__ void svd_gather2_k(const double* __restrict__ V, long ldv, int q,
                              const int* __restrict__ perm, double* __restrict__ out, long s0,
                              long s1)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= q) return;
    out[(long)j * s0 + (long)i * s1] = V[(long)perm[j] * ldv + i];
}
hipError_t svd_gather(const double* X, long ldx, int p, const double* V, long ldv, int q,
                      const int* perm, const double* scl, int k, double* out1, long s10, long s11,
                      double* out2, long s20, long s21, hipStream_t s)
{
    if (k <= 0) return hipSuccess;
    dim3 g1((p + 255) / 256, k), g2((q + 255) / 256, k);
    svd_gather1_k<<<g1, 256, 0, s>>>(X, ldx, p, perm, scl, out1, s10, s11);
    svd_gather2_k<<<g2, 256, 0, s>>>(V, ldv, q, perm, out2, s20, s21);
    return hipGetLastError();
}

__global__ void set_unit_col_k(double* x, int p, int pos)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < p; i += gridDim.x * 256) x[i] = (i == pos) ? 1.0 : 0.0;
}
hipError_t set_unit_col(double* x, int p, int pos, hipStream_t s)
{
    int blocks = (p + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    set_unit_col_k<<<blocks, 256, 0, s>>>(x, p, pos);
    return hipGetLastError();
}

__global__ void scale_cols_k(double* A, long lda, int r, int c, const double* sv)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (j < c && i < r) A[(long)i * lda + j] *= sv[j];
}
__global__ void scale_rows_k(double* A, long lda, int r, int c, const double* sv)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (j < c && i < r) A[(long)i * lda + j] *= sv[i];
}
hipError_t scale_cols(double* A, long lda, int r, int c, const double* sv, hipStream_t s)
{
    if (r <= 0 || c <= 0) return hipSuccess;
    dim3 grid((c + 255) / 256, r);
    scale_cols_k<<<grid, 256, 0, s>>>(A, lda, r, c, sv);
    return hipGetLastError();
}
hipError_t scale_rows(double* A, long lda, int r, int c, const double* sv, hipStream_t s)
{
    if (r <= 0 || c <= 0) return hipSuccess;
    dim3 grid((c + 255) / 256, r);
    scale_rows_k<<<grid, 256, 0, s>>>(A, lda, r, c, sv);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// CholeskyQR pieces.  chol_diag_k factors one CB x CB diagonal block of a
// symmetric positive definite G and writes L_kk (lower) and its inverse; a
// non-positive or non-finite pivot sets flag[0] = 1.
#define CB 64
#define CW 16  // columns of the trailing matrix (rows of the inverse) per wave
// v from lane `lane` (wave-uniform index) of the wavefront, as a register
// broadcast (two 32-bit readlanes)
__device__ __forceinline__ double wave_bcast(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Four wavefronts, lane r = row r.  Wave w keeps columns 16w .. 16w+15 of the
// not yet factored matrix, a[kk] = A[r][16w + kk], and rows 16w .. 16w+15 of
// the right-hand side of L X = I, b[kk] = B[16w + kk][r] (lane r = column r
// of X).  Right-looking step j, owned by wave j / 16 (register index j % 16 a
// compile-time constant: the 16 steps of a column block are unrolled):
//   owner: pivot A[j][j] from lane j, L[r][j] = A[r][j] / L[j][j] (the whole
//          column, zero above the diagonal) and X[j][r] = B[j][r] / L[j][j]
//          into LDS; X[j][:] is row j of L^-1;
//   all:   A[r][k] -= L[r][j] L[k][j] on their columns k and
//          B[i][r] -= L[i][j] X[j][r] on their rows i, L[k][j] read from LDS
//          as a broadcast.
// Entries above the diagonal of column j of L are zero, so the full-width
// updates need no guard; columns and rows already consumed take values that
// are never read again.  The column and row buffers alternate between two
// LDS slots by step parity, so one barrier per step orders them: a wave can
// only overwrite a slot after every wave has passed the next barrier, that
// is, after every wave has read it.
__device__ __forceinline__ void chol_block(const double* __restrict__ G, long ldg, int k0, int nbk,
                                           double* __restrict__ L, long ldl, double* __restrict__ Linv, long ldi,
                                           int* __restrict__ flag, double (*Ls)[CB + 1], double (*col)[CB],
                                           double (*xrow)[CB])
{
    const int tid = threadIdx.x, r = tid & 63, wid = tid >> 6;
    // coalesced load (lane = column, one row per wave per round); a block
    // narrower than CB (the last one) is completed with the identity
#pragma unroll
    for (int i0 = 0; i0 < CB; i0 += 4 * 4) {
        double tmp[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * 4 + wid;
            tmp[u] = (i < nbk && r < nbk) ? G[(long)(k0 + i) * ldg + k0 + r] : (i == r ? 1.0 : 0.0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) Ls[i0 + u * 4 + wid][r] = tmp[u];
    }
    __syncthreads();
    double a[CW], bv[CW];
#pragma unroll
    for (int kk = 0; kk < CW; kk++) {
        a[kk] = Ls[r][wid * CW + kk];
        bv[kk] = (wid * CW + kk == r) ? 1.0 : 0.0;
    }
    __syncthreads();  // every row is in registers before Ls takes L
    int bad = 0;
    for (int jb = 0; jb < CB / CW; jb++) {
#pragma unroll
        for (int jj = 0; jj < CW; jj++) {
            const int j = jb * CW + jj, slot = jj & 1;
            if (wid == jb) {
                const double piv = wave_bcast(a[jj], j);
                const bool ok = piv > 0.0 && isfinite(piv);
                bad |= ok ? 0 : 1;
                const double d = ok ? sqrt(piv) : 1.0;
                const double inv = 1.0 / d;
                const double cj = r > j ? a[jj] * inv : (r == j ? d : 0.0);
                const double xj = bv[jj] * inv;
                col[slot][r] = cj;
                xrow[slot][r] = xj;
                Ls[r][j] = cj;
                if (j < nbk && r < nbk) Linv[(long)(k0 + j) * ldi + k0 + r] = xj;
            }
            __syncthreads();
            const double lr = col[slot][r];   // L[r][j]
            const double xr = xrow[slot][r];  // X[j][r]
#pragma unroll
            for (int kk = 0; kk < CW; kk += 2) {
                const double2 lk = *(const double2*)&col[slot][wid * CW + kk];  // L[k][j], L[k+1][j]
                a[kk] = fma(-lr, lk.x, a[kk]);
                a[kk + 1] = fma(-lr, lk.y, a[kk + 1]);
                bv[kk] = fma(-lk.x, xr, bv[kk]);
                bv[kk + 1] = fma(-lk.y, xr, bv[kk + 1]);
            }
        }
    }
    __syncthreads();
    // L (lower; Ls holds zeros above the diagonal), coalesced rows
    for (int i = wid; i < nbk; i += 4)
        if (r < nbk) L[(long)(k0 + i) * ldl + k0 + r] = Ls[i][r];
    if (r == 0 && bad) flag[0] = 1;
}

__global__ void __launch_bounds__(256) chol_diag_k(const double* __restrict__ G, long ldg, int k0, int nbk,
                                                   double* __restrict__ L, long ldl,
                                                   double* __restrict__ Linv, long ldi,
                                                   int* __restrict__ flag)
{
    __shared__ double Ls[CB][CB + 1];
    __shared__ __attribute__((aligned(16))) double col[2][CB];
    __shared__ double xrow[2][CB];
    chol_block(G, ldg, k0, nbk, L, ldl, Linv, ldi, flag, Ls, col, xrow);
}

// one diagonal block per workgroup, one job per chain
__global__ void __launch_bounds__(256) chol_diag_batch_k(const CholJob* __restrict__ jobs)
{
    __shared__ double Ls[CB][CB + 1];
    __shared__ __attribute__((aligned(16))) double col[2][CB];
    __shared__ double xrow[2][CB];
    const CholJob& j = jobs[blockIdx.x];
    chol_block(j.G, j.ldg, j.k0, j.nbk, j.L, j.ldl, j.Linv, j.ldi, j.flag, Ls, col, xrow);
}

hipError_t chol_diag_batch(const CholJob* jobs, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    chol_diag_batch_k<<<n, 256, 0, s>>>(jobs);
    return hipGetLastError();
}

hipError_t chol_diag(const double* G, long ldg, int k0, int nbk, double* L, long ldl, double* Linv, long ldi,
                     int* flag, hipStream_t s)
{
    if (nbk <= 0 || nbk > CB) return hipErrorInvalidValue;
    chol_diag_k<<<1, 256, 0, s>>>(G, ldg, k0, nbk, L, ldl, Linv, ldi, flag);
    return hipGetLastError();
}

// E = G - I on the leading nl x nl block, zero elsewhere (n x n, ld n)
__global__ void __launch_bounds__(256) eye_sub_k(const double* __restrict__ G, int n, int nl, double* __restrict__ E)
{
    const long nn = (long)n * n;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < nn; e += (long)gridDim.x * 256) {
        const int i = (int)(e / n), j = (int)(e - (long)i * n);
        E[e] = (i < nl && j < nl) ? G[e] - (i == j ? 1.0 : 0.0) : 0.0;
    }
}
hipError_t eye_sub(const double* G, int n, int nl, double* E, hipStream_t s)
{
    long blocks = ((long)n * n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    eye_sub_k<<<blocks, 256, 0, s>>>(G, n, nl, E);
    return hipGetLastError();
}

// X = I + c1 E + c2 E2 + c3 E3 and Y = I + d1 E + d2 E2 + d3 E3 (n x n)
__global__ void __launch_bounds__(256) series3_k(const double* __restrict__ E, const double* __restrict__ E2,
                                                 const double* __restrict__ E3, int n, double c1, double c2,
                                                 double c3, double d1, double d2, double d3,
                                                 double* __restrict__ X, double* __restrict__ Y)
{
    const long nn = (long)n * n;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < nn; e += (long)gridDim.x * 256) {
        const int i = (int)(e / n), j = (int)(e - (long)i * n);
        const double id = i == j ? 1.0 : 0.0, a = E[e], b = E2[e], c = E3[e];
        X[e] = id + c1 * a + c2 * b + c3 * c;
        Y[e] = id + d1 * a + d2 * b + d3 * c;
    }
}
hipError_t series3(const double* E, const double* E2, const double* E3, int n, const double c[3], const double d[3],
                   double* X, double* Y, hipStream_t s)
{
    long blocks = ((long)n * n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    series3_k<<<blocks, 256, 0, s>>>(E, E2, E3, n, c[0], c[1], c[2], d[0], d[1], d[2], X, Y);
    return hipGetLastError();
}

// G[i][i] += factor * trace(G[0:n, 0:n]) for i < n  (the shift of shifted
// CholeskyQR; trace(A^T A) = ||A||_F^2 bounds ||A||_2^2 from above)
__global__ void __launch_bounds__(256) shift_diag_k(double* __restrict__ G, long ldg, int n, double factor)
{
    __shared__ double red[4];
    __shared__ double sh;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc += G[(long)i * ldg + i];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) sh = factor * (red[0] + red[1] + red[2] + red[3]);
    __syncthreads();
    const double v = sh;
    for (int i = threadIdx.x; i < n; i += 256) G[(long)i * ldg + i] += v;
}
hipError_t shift_diag(double* G, long ldg, int n, double factor, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    shift_diag_k<<<1, 256, 0, s>>>(G, ldg, n, factor);
    return hipGetLastError();
}

// G[i*ldg + i] += v for i in [i0, i1)
__global__ void add_diag_k(double* G, long ldg, int i0, int i1, double v)
{
    const int i = i0 + blockIdx.x * 256 + threadIdx.x;
    if (i < i1) G[(long)i * ldg + i] += v;
}
hipError_t add_diag(double* G, long ldg, int i0, int i1, double v, hipStream_t s)
{
    if (i1 <= i0) return hipSuccess;
    add_diag_k<<<(i1 - i0 + 255) / 256, 256, 0, s>>>(G, ldg, i0, i1, v);
    return hipGetLastError();
}

// out[0] = max_{i,j<n} |G[i][j] - delta_ij|   (deviation from the identity)
__global__ void __launch_bounds__(256) eye_dev_k(const double* __restrict__ G, long ldg, int n,
                                                 double* __restrict__ part)
{
    __shared__ double red[4];
    double m = 0.0;
    const long tot = (long)n * n;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
        const long i = e / n, j = e - i * n;
        const double v = G[i * ldg + j] - (i == j ? 1.0 : 0.0);
        m = fmax(m, isfinite(v) ? fabs(v) : 1e300);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}
hipError_t eye_deviation(const double* G, long ldg, int n, double* out, double* part, hipStream_t s)
{
    long G_ = ((long)n * n + 255) / 256;
    if (G_ > 1024) G_ = 1024;
    eye_dev_k<<<G_, 256, 0, s>>>(G, ldg, n, part);
    absmax2_k<<<1, 256, 0, s>>>(part, (int)G_, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched vector kernels: one job per problem (chain), problem index = last
// grid dimension.
#define BV_EPT 8
#define BV_BLOCK 256
#define BV_CHUNK (BV_BLOCK * BV_EPT)

// stage 1: part[(q * kmax + i) * G + b] = dot of chunk b of V_i and w
__global__ void __launch_bounds__(BV_BLOCK) bmdot1_k(const DotJob* __restrict__ jobs, int kmax, int G,
                                                     double* __restrict__ part)
{
    __shared__ double red[BV_BLOCK / 64];
    const DotJob& j = jobs[blockIdx.y];
    const int b = blockIdx.x;
    double* pq = part + (size_t)blockIdx.y * kmax * G;
    const long base = (long)b * BV_CHUNK + threadIdx.x;
    double wr[BV_EPT];
#pragma unroll
    for (int e = 0; e < BV_EPT; e++) {
        const long idx = base + (long)e * BV_BLOCK;
        wr[e] = idx < j.n ? j.w[idx] : 0.0;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = 0; i < j.k; i++) {
        const double* vi = j.V + (long)i * j.ldv;
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < BV_EPT; e++) {
            const long idx = base + (long)e * BV_BLOCK;
            if (idx < j.n) acc += vi[idx] * wr[e];
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) red[wid] = acc;
        __syncthreads();
        if (threadIdx.x == 0) pq[(size_t)i * G + b] = (red[0] + red[1]) + (red[2] + red[3]);
        __syncthreads();
    }
}
// stage 2: out[i] = sum_b part, one workgroup per (problem, row)
__global__ void __launch_bounds__(256) bmdot2_k(const DotJob* __restrict__ jobs, int kmax, int G,
                                                const double* __restrict__ part)
{
    __shared__ double red[4];
    const DotJob& j = jobs[blockIdx.y];
    const int i = blockIdx.x;
    if (i >= j.k) return;
    const long nb = (j.n + BV_CHUNK - 1) / BV_CHUNK;
    const double* pq = part + ((size_t)blockIdx.y * kmax + i) * G;
    double acc = 0.0;
    for (long b = threadIdx.x; b < nb; b += 256) acc += pq[b];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) j.out[i] = (red[0] + red[1]) + (red[2] + red[3]);
}
hipError_t bmdot(const DotJob* jobs, int n, int kmax, long nmax, double* part, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int G = (int)((nmax + BV_CHUNK - 1) / BV_CHUNK);
    bmdot1_k<<<dim3(G, n), BV_BLOCK, 0, s>>>(jobs, kmax, G, part);
    bmdot2_k<<<dim3(kmax, n), 256, 0, s>>>(jobs, kmax, G, part);
    return hipGetLastError();
}
long bmdot_part_size(int n, int kmax, long nmax) { return (long)n * kmax * ((nmax + BV_CHUNK - 1) / BV_CHUNK); }

// Fused Lanczos step, batched: block (b, q) owns chunk b of job q's vectors
// (BV_CHUNK entries as BV_EPT / 2 pairs per thread), the layout of lz_step_dots_k.
#define BV_PAIRS (BV_EPT / 2)
__global__ void __launch_bounds__(BV_BLOCK) bstep_dots_k(const StepJob* __restrict__ jobs, int kmax, int G,
                                                         double* __restrict__ part)
{
    __shared__ double red[MD_MAXK][BV_BLOCK / 64];
    const StepJob& j = jobs[blockIdx.y];
    const long c0 = (long)blockIdx.x * BV_CHUNK;
    if (c0 >= j.n) return;  // the finalising sum reads only this job's chunks
    const double a = j.alpha[0];
    const double beta = j.beta_sq ? sqrt(fmax(*j.beta_sq, 0.0)) : j.beta;
    const double* vj = j.V + (long)(j.k - 1) * j.ldv;
    const double* vm = j.k >= 2 ? j.V + (long)(j.k - 2) * j.ldv : nullptr;
    double2 wr[BV_PAIRS];
#pragma unroll
    for (int e = 0; e < BV_PAIRS; e++) {
        const long idx = c0 + 2 * ((long)e * BV_BLOCK + threadIdx.x);
        if (idx < j.n) {
            double2 x = *(const double2*)(j.w + idx);
            const double2 p = *(const double2*)(vj + idx);
            x.x -= a * p.x;
            x.y -= a * p.y;
            if (vm) {
                const double2 q = *(const double2*)(vm + idx);
                x.x -= beta * q.x;
                x.y -= beta * q.y;
            }
            *(double2*)(j.w + idx) = x;
            wr[e] = x;
        } else {
            wr[e] = make_double2(0.0, 0.0);
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int i = 0; i < j.k; i++) {
        const double* vi = j.V + (long)i * j.ldv;
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < BV_PAIRS; e++) {
            const long idx = c0 + 2 * ((long)e * BV_BLOCK + threadIdx.x);
            if (idx < j.n) {
                const double2 p = *(const double2*)(vi + idx);
                acc += p.x * wr[e].x + p.y * wr[e].y;
            }
        }
        acc = wave_sum(acc);
        if (lane == 0) red[i][wid] = acc;
    }
    __syncthreads();
    double* pq = part + (size_t)blockIdx.y * kmax * G;
    for (int i = threadIdx.x; i < j.k; i += BV_BLOCK) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < BV_BLOCK / 64; q++) sum += red[i][q];
        pq[(size_t)i * G + blockIdx.x] = sum;
    }
}
// out[i] = sum over the job's chunks, fixed order; one workgroup per (row, job)
__global__ void __launch_bounds__(256) bstep_sum_k(const StepJob* __restrict__ jobs, int kmax, int G,
                                                   const double* __restrict__ part)
{
    __shared__ double red[4];
    const StepJob& j = jobs[blockIdx.y];
    const int i = blockIdx.x;
    if (i >= j.k) return;
    const long nb = (j.n + BV_CHUNK - 1) / BV_CHUNK;
    const double* pq = part + ((size_t)blockIdx.y * kmax + i) * G;
    double acc = 0.0;
    for (long b = threadIdx.x; b < nb; b += 256) acc += pq[b];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) j.out[i] = (red[0] + red[1]) + (red[2] + red[3]);
}
hipError_t bstep_dots(const StepJob* jobs, int n, int kmax, long nmax, double* part, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (kmax < 1 || kmax > MD_MAXK) return hipErrorInvalidValue;
    const int G = (int)((nmax + BV_CHUNK - 1) / BV_CHUNK);
    bstep_dots_k<<<dim3(G, n), BV_BLOCK, 0, s>>>(jobs, kmax, G, part);
    bstep_sum_k<<<dim3(kmax, n), 256, 0, s>>>(jobs, kmax, G, part);
    return hipGetLastError();
}

// chunks in reverse order of bstep_dots_k (see lz_update_norm_k)
__global__ void __launch_bounds__(BV_BLOCK) bupdate_norm_k(const UpdJob* __restrict__ jobs, int G,
                                                           double* __restrict__ part)
{
    __shared__ double c[MD_MAXK];
    __shared__ double red[BV_BLOCK / 64];
    const UpdJob& j = jobs[blockIdx.y];
    const int nb = (int)((j.n + BV_CHUNK - 1) / BV_CHUNK);
    if ((int)blockIdx.x >= nb) return;
    for (int i = threadIdx.x; i < j.k; i += BV_BLOCK) c[i] = j.coef[i];
    __syncthreads();
    const int blk = nb - 1 - blockIdx.x;
    const long c0 = (long)blk * BV_CHUNK;
    double2 acc[BV_PAIRS];
#pragma unroll
    for (int e = 0; e < BV_PAIRS; e++) {
        const long idx = c0 + 2 * ((long)e * BV_BLOCK + threadIdx.x);
        acc[e] = idx < j.n ? *(const double2*)(j.w + idx) : make_double2(0.0, 0.0);
    }
    for (int i = 0; i < j.k; i++) {
        const double* vi = j.V + (long)i * j.ldv;
        const double ci = c[i];
#pragma unroll
        for (int e = 0; e < BV_PAIRS; e++) {
            const long idx = c0 + 2 * ((long)e * BV_BLOCK + threadIdx.x);
            if (idx < j.n) {
                const double2 p = *(const double2*)(vi + idx);
                acc[e].x -= ci * p.x;
                acc[e].y -= ci * p.y;
            }
        }
    }
    double sq = 0.0;
#pragma unroll
    for (int e = 0; e < BV_PAIRS; e++) {
        const long idx = c0 + 2 * ((long)e * BV_BLOCK + threadIdx.x);
        if (idx < j.n) {
            *(double2*)(j.w + idx) = acc[e];
            sq += acc[e].x * acc[e].x + acc[e].y * acc[e].y;
        }
    }
    sq = wave_sum(sq);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * G + blk] = (red[0] + red[1]) + (red[2] + red[3]);
}
__global__ void __launch_bounds__(256) bupdate_sum_k(const UpdJob* __restrict__ jobs, int G,
                                                     const double* __restrict__ part)
{
    __shared__ double red[4];
    const UpdJob& j = jobs[blockIdx.x];
    const long nb = (j.n + BV_CHUNK - 1) / BV_CHUNK;
    const double* pq = part + (size_t)blockIdx.x * G;
    double acc = 0.0;
    for (long b = threadIdx.x; b < nb; b += 256) acc += pq[b];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) j.out[0] = (red[0] + red[1]) + (red[2] + red[3]);
}
hipError_t bupdate_norm(const UpdJob* jobs, int n, long nmax, double* part, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int G = (int)((nmax + BV_CHUNK - 1) / BV_CHUNK);
    bupdate_norm_k<<<dim3(G, n), BV_BLOCK, 0, s>>>(jobs, G, part);
    bupdate_sum_k<<<n, 256, 0, s>>>(jobs, G, part);
    return hipGetLastError();
}

// deviation from the identity, many matrices: 64 workgroups per job
__global__ void __launch_bounds__(256) beye_dev_k(const EyeJob* __restrict__ jobs, double* __restrict__ part)
{
    __shared__ double red[4];
    const EyeJob& j = jobs[blockIdx.y];
    double m = 0.0;
    const long tot = (long)j.n * j.n;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < tot; e += (long)gridDim.x * 256) {
        const long i = e / j.n, c = e - i * j.n;
        const double v = j.G[i * j.ld + c] - (i == c ? 1.0 : 0.0);
        m = fmax(m, isfinite(v) ? fabs(v) : 1e300);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}
__global__ void __launch_bounds__(64) beye_max_k(const EyeJob* __restrict__ jobs, int G, const double* __restrict__ part)
{
    double m = 0.0;
    for (int b = threadIdx.x; b < G; b += 64) m = fmax(m, part[(size_t)blockIdx.x * G + b]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (threadIdx.x == 0) jobs[blockIdx.x].out[0] = m;
}
hipError_t beye_deviation(const EyeJob* jobs, int n, double* part, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    beye_dev_k<<<dim3(64, n), 256, 0, s>>>(jobs, part);
    beye_max_k<<<n, 64, 0, s>>>(jobs, 64, part);
    return hipGetLastError();
}

// strided 2-D copies, one job per problem: grid (column blocks, rows, problem)
__global__ void __launch_bounds__(256) bcopy2_k(const Copy2Job* __restrict__ jobs)
{
    const Copy2Job& j = jobs[blockIdx.z];
    const int col = blockIdx.x * 256 + threadIdx.x;
    const int row = blockIdx.y;
    if (col >= j.co || row >= j.ro) return;
    j.out[row * j.so0 + col * j.so1] = (row < j.r && col < j.c) ? j.in[row * j.si0 + col * j.si1] : 0.0;
}
hipError_t bcopy2(const Copy2Job* jobs, int n, int romax, int comax, hipStream_t s)
{
    if (n <= 0 || romax <= 0 || comax <= 0) return hipSuccess;
    bcopy2_k<<<dim3((comax + 255) / 256, romax, n), 256, 0, s>>>(jobs);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) badd_diag_k(const DiagJob* __restrict__ jobs)
{
    const DiagJob& j = jobs[blockIdx.y];
    const int i = j.i0 + blockIdx.x * 256 + threadIdx.x;
    if (i < j.i1) j.G[(long)i * j.ld + i] += j.v;
}
hipError_t badd_diag(const DiagJob* jobs, int n, int lenmax, hipStream_t s)
{
    if (n <= 0 || lenmax <= 0) return hipSuccess;
    badd_diag_k<<<dim3((lenmax + 255) / 256, n), 256, 0, s>>>(jobs);
    return hipGetLastError();
}

// w = beta w + cscale sum_i coef[i] V_i
__global__ void __launch_bounds__(256) bgemv_k(const GemvJob* __restrict__ jobs)
{
    __shared__ double cf[MD_MAXK];
    const GemvJob& j = jobs[blockIdx.y];
    for (int i = threadIdx.x; i < j.k; i += 256) cf[i] = j.coef[i] * j.cscale;
    __syncthreads();
    for (long idx = blockIdx.x * 256L + threadIdx.x; idx < j.n; idx += (long)gridDim.x * 256) {
        double acc = (j.beta == 0.0) ? 0.0 : j.beta * j.w[idx];
        for (int i = 0; i < j.k; i++) acc += cf[i] * j.V[(long)i * j.ldv + idx];
        j.w[idx] = acc;
    }
}
hipError_t bgemv(const GemvJob* jobs, int n, long nmax, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (nmax + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    bgemv_k<<<dim3(blocks, n), 256, 0, s>>>(jobs);
    return hipGetLastError();
}

// w -= alpha[0] v + beta u
__global__ void __launch_bounds__(256) bthree_k(const ThreeJob* __restrict__ jobs)
{
    const ThreeJob& j = jobs[blockIdx.y];
    const double a = j.alpha[0];
    for (long i = blockIdx.x * 256L + threadIdx.x; i < j.n; i += (long)gridDim.x * 256) {
        double r = j.w[i] - a * j.v[i];
        if (j.u) r -= j.beta * j.u[i];
        j.w[i] = r;
    }
}
hipError_t bthree(const ThreeJob* jobs, int n, long nmax, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (nmax + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    bthree_k<<<dim3(blocks, n), 256, 0, s>>>(jobs);
    return hipGetLastError();
}

// y = a x
__global__ void __launch_bounds__(256) bscale_k(const ScaleJob* __restrict__ jobs)
{
    const ScaleJob& j = jobs[blockIdx.y];
    const double a = j.a_sq ? 1.0 / sqrt(*j.a_sq) : j.a;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < j.n; i += (long)gridDim.x * 256) j.y[i] = a * j.x[i];
}
hipError_t bscale(const ScaleJob* jobs, int n, long nmax, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (nmax + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    bscale_k<<<dim3(blocks, n), 256, 0, s>>>(jobs);
    return hipGetLastError();
}

// the scaling of the next Krylov vector and its FP32 copy in one pass
__global__ void __launch_bounds__(256) bscale_conv_k(const ScaleConvJob* __restrict__ jobs)
{
    const ScaleConvJob& j = jobs[blockIdx.y];
    const double a = j.a_sq ? 1.0 / sqrt(*j.a_sq) : j.a;
    const double* x = j.x;  // y may be x (in place)
    double* y = j.y;
    float* __restrict__ y32 = j.y32;
    const long n = j.n;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double v = a * x[i];
        y[i] = v;
        y32[i] = (float)v;
    }
}
hipError_t bscale_conv(const ScaleConvJob* jobs, int n, long nmax, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (nmax + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    bscale_conv_k<<<dim3(blocks, n), 256, 0, s>>>(jobs);
    return hipGetLastError();
}

// FP64 -> FP32 copy
__global__ void __launch_bounds__(256) bconvert_k(const ConvJob* __restrict__ jobs)
{
    const ConvJob& j = jobs[blockIdx.y];
    for (long i = blockIdx.x * 256L + threadIdx.x; i < j.n; i += (long)gridDim.x * 256) j.y[i] = (float)j.x[i];
}
hipError_t bconvert(const ConvJob* jobs, int n, long nmax, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (nmax + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    bconvert_k<<<dim3(blocks, n), 256, 0, s>>>(jobs);
    return hipGetLastError();
}

// MPO mixing (FP32 in, TO out) for many problems: grid (Bn blocks, X, problem).
// The workgroup's mixing table (one problem's) is staged in LDS first when it
// fits: read from global memory, every entry would be a chain of dependent
// loads (offset, value, selector, then the data) per output.
#define BW_MAXO 64
#define BW_MAXE 512
template <typename TO>
__global__ void __launch_bounds__(256) bwmix_k(const MixJob* __restrict__ jobs)
{
    __shared__ int s_optr[BW_MAXO + 1];
    __shared__ long s_ooff[BW_MAXO];
    __shared__ int s_osel[BW_MAXO];
    __shared__ long s_eoff[BW_MAXE];
    __shared__ float s_eval[BW_MAXE];
    __shared__ int s_esel[BW_MAXE];
    const MixJob& j = jobs[blockIdx.z];
    const long x = blockIdx.y;
    if (x >= j.X) return;  // the whole workgroup
    const MixTable& t = j.t;
    const int nout = t.nout, nnz = t.nnz;
    const bool idt = t.esel != nullptr;
    const bool staged = nout <= BW_MAXO && nnz <= BW_MAXE;
    if (staged) {
        for (int i = threadIdx.x; i <= nout; i += 256) s_optr[i] = t.optr[i];
        for (int i = threadIdx.x; i < nout; i += 256) {
            s_ooff[i] = t.ooff[i];
            s_osel[i] = idt ? t.osel[i] : 0;
        }
        for (int i = threadIdx.x; i < nnz; i += 256) {
            s_eoff[i] = t.eoff[i];
            s_eval[i] = t.evalf[i];
            s_esel[i] = idt ? t.esel[i] : 0;
        }
        __syncthreads();
    }
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= j.Bn) return;
    const float* ip = (const float*)j.in + x * j.isx + b;
    TO* op = (TO*)j.out + x * j.osx + b;
    const float* ip2 = idt ? j.in2 + x * j.isx2 + b : nullptr;
    if (staged) {
        for (int o = 0; o < nout; o++) {
            const int e0 = s_optr[o], e1 = s_optr[o + 1];
            float acc = 0.0f;
#pragma unroll 4
            for (int e = e0; e < e1; e++) acc += s_eval[e] * (s_esel[e] ? ip2[s_eoff[e]] : ip[s_eoff[e]]);
            if (s_osel[o]) j.out2[x * j.osx2 + s_ooff[o] + b] = (double)acc;
            else op[s_ooff[o]] = cvt_from_f<TO>(acc);
        }
        return;
    }
    for (int o = 0; o < nout; o++) {
        const int e0 = t.optr[o], e1 = t.optr[o + 1];
        float acc = 0.0f;
        for (int e = e0; e < e1; e++) acc += t.evalf[e] * (idt && t.esel[e] ? ip2[t.eoff[e]] : ip[t.eoff[e]]);
        if (idt && t.osel[o]) j.out2[x * j.osx2 + t.ooff[o] + b] = (double)acc;
        else op[t.ooff[o]] = cvt_from_f<TO>(acc);
    }
}
hipError_t bwmix(int tout, const MixJob* jobs, int n, int Xmax, int Bnmax, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    dim3 grid((Bnmax + 255) / 256, Xmax, n);
    if (tout == T_F32) bwmix_k<float><<<grid, 256, 0, s>>>(jobs);
    else if (tout == T_F64) bwmix_k<double><<<grid, 256, 0, s>>>(jobs);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace tn
