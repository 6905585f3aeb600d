"""GPU parity of the hot-path kernels against the CPU oracle (and numpy).

Every case calls the HIP path through the C ABI (include/tn_abi.h).
Tolerances: FP64 1e-13 relative (bit-level agreement is not expected since
summation orders differ); FP32/BF16 operands are compared against fp64
products of the *rounded* operands, which pins the matrix-core lane layout.
"""
import numpy as np
import pytest

import oracle as O
import tensornetworks_amd as tn

pytestmark = pytest.mark.gpu


def rand(shape, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, shape)


def round_fp32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def round_bf16(x):
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


ROUND = {0: lambda x: np.asarray(x, dtype=np.float64), 1: round_fp32, 2: round_bf16}
LAYOUTS = [(True, True), (True, False), (False, False), (False, True)]


@pytest.mark.parametrize("akm,bkm", LAYOUTS)
@pytest.mark.parametrize("tin,tout", [(0, 0), (1, 1), (1, 0), (2, 1), (2, 0), (2, 2)])
def test_gemm_matches_rounded_fp64(ctx, tin, tout, akm, bkm):
    M, N, K, batch = 96, 160, 128, 2
    A = rand((batch, M, K), 1)
    B = rand((batch, K, N), 2)
    Ar, Br = ROUND[tin](A), ROUND[tin](B)
    ref = Ar @ Br
    Ast = A if akm else np.ascontiguousarray(A.transpose(0, 2, 1))
    Bst = np.ascontiguousarray(B.transpose(0, 2, 1)) if bkm else B
    C = ctx.gemm_test(Ast, Bst, tin, tout, akm, bkm, M, N, K, batch)
    scale = np.abs(Ar).max() * np.abs(Br).max() * K
    err = np.abs(C - ref).max() / scale
    # accumulation: fp64 ~1e-16, fp32 ~ K*2^-24; output rounding for bf16 out
    tol = {0: 1e-15, 1: 2e-6}.get(tin, 2e-6)
    if tout == 2:
        tol = 8e-3 / np.sqrt(K)
    assert err < tol, (tin, tout, akm, bkm, err)


@pytest.mark.parametrize("tile", range(1, 11))
@pytest.mark.parametrize("tin", [0, 1, 2])
def test_gemm_every_tile(ctx, tin, tile):
    M, N, K = 128, 128, 256
    A = rand((1, M, K), 3)
    B = rand((1, K, N), 4)
    ref = ROUND[tin](A) @ ROUND[tin](B)
    C = ctx.gemm_test(A, np.ascontiguousarray(B.transpose(0, 2, 1)), tin, 0, True, True, M, N, K, 1, tile)
    err = np.abs(C - ref).max() / (K * 1.0)
    assert err < (1e-15 if tin == 0 else 2e-6), (tile, err)


@pytest.mark.parametrize("ks", [2, 4])
@pytest.mark.parametrize("tin,tout", [(0, 0), (1, 1), (1, 0), (2, 1), (2, 0)])
@pytest.mark.parametrize("akm,bkm", [(True, True), (True, False), (False, False)])
def test_gemm_split_k(ctx, tin, tout, akm, bkm, ks):
    # K split into ks partial products reduced by a second kernel: the same
    # product as the plain GEMM to the operand precision
    M, N, K = 96, 160, 512
    A = rand((1, M, K), 5)
    B = rand((1, K, N), 6)
    ref = ROUND[tin](A) @ ROUND[tin](B)
    Ain = A if akm else np.ascontiguousarray(A.transpose(0, 2, 1))
    Bin = np.ascontiguousarray(B.transpose(0, 2, 1)) if bkm else B
    C = ctx.gemm_test(Ain, Bin, tin, tout, akm, bkm, M, N, K, 1, 0, ks)
    # summation-order rounding: ~sqrt(K) u |C|, |C| ~ sqrt(K)/3
    err = np.abs(C - ref).max() / np.sqrt(K)
    assert err < (1e-14 if tin == 0 else 2e-6), (tin, tout, ks, err)
    # accumulate=0 semantics: a second call gives the same result
    C2 = ctx.gemm_test(Ain, Bin, tin, tout, akm, bkm, M, N, K, 1, 0, ks)
    assert np.array_equal(C, C2)


def sym_env(Dw, D, seed):
    """Environment whose every w-slice is symmetric (reference layout)."""
    F = rand((Dw, D, D), seed)
    F = 0.5 * (F + F.transpose(0, 2, 1))
    # reference layout F[w + Dw*(i + D*j)]
    return np.ascontiguousarray(F.transpose(2, 1, 0)).ravel()


def sym_W(Dw1, Dw2, d, seed):
    W = rand((Dw1, Dw2, d, d), seed)
    W = 0.5 * (W + W.transpose(0, 1, 3, 2))
    W[np.abs(W) < 0.3] = 0.0  # sparsity like a real MPO
    return np.ascontiguousarray(W)


HEFF_CASES = [(2, 1, 2, 1, 5), (2, 8, 8, 5, 5), (3, 27, 40, 5, 5), (2, 37, 19, 5, 5), (3, 64, 64, 5, 5),
              (2, 16, 1, 5, 1), (4, 5, 9, 3, 4)]


@pytest.mark.parametrize("d,D1,D2,Dw1,Dw2", HEFF_CASES)
def test_heff_apply_fp64_matches_oracle(ctx, d, D1, D2, Dw1, Dw2):
    W = sym_W(Dw1, Dw2, d, 10)
    L = rand(Dw1 * D1 * D1, 11)
    R = rand(Dw2 * D2 * D2, 12)
    x = rand(d * D1 * D2, 13)
    y_ref = O.heff_apply_ref(W, L, R, x, d, D1, D2)
    ctx.site_bind(W, L, R, d, D1, D2, tn.TN_FP64)
    y = ctx.heff_apply(x)
    err = np.abs(y - y_ref).max() / np.abs(y_ref).max()
    assert err < 1e-13, err
    if d * D1 * D2 <= 400:
        H = O.heff_dense(W, L, R, d, D1, D2)
        assert np.abs(H @ x - y).max() / np.abs(y).max() < 1e-13


@pytest.mark.parametrize("prec,tol", [(tn.TN_FP32, 5e-6), (tn.TN_BF16, 2e-2)])
def test_heff_apply_reduced_precision(ctx, prec, tol):
    d, D1, D2 = 3, 64, 96
    W = sym_W(5, 5, d, 20)
    L, R, x = rand(5 * D1 * D1, 21), rand(5 * D2 * D2, 22), rand(d * D1 * D2, 23)
    y_ref = O.heff_apply_ref(W, L, R, x, d, D1, D2)
    ctx.site_bind(W, L, R, d, D1, D2, prec)
    y = ctx.heff_apply(x)
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    assert err < tol, err


def with_identity_channel(F, Dw, D, w):
    """Environment (reference layout F[w + Dw*(i + D*j)]) with channel w set to the identity."""
    Fr = F.reshape(D, D, Dw).copy()
    Fr[:, :, w] = np.eye(D)
    return Fr.ravel()


# (d, D1, D2, Dw1, Dw2, identity channel of L, of R): first/last channels,
# one side only, padded (D not a multiple of 32) and unpadded shapes
IDENT_CASES = [(2, 8, 8, 5, 5, 0, 4), (3, 27, 40, 5, 5, 4, 0), (2, 37, 19, 5, 5, 0, -1), (3, 64, 64, 5, 5, -1, 4),
               (4, 5, 9, 3, 4, 2, 3), (2, 32, 33, 2, 2, 1, 0)]


@pytest.mark.parametrize("d,D1,D2,Dw1,Dw2,wl,wr", IDENT_CASES)
@pytest.mark.parametrize("prec,tol", [(tn.TN_FP64, 1e-13), (tn.TN_FP32, 5e-6), (tn.TN_BF16, 2e-2),
                                      (tn.TN_BF16X3, 5e-6)])
def test_heff_apply_identity_channels_match_oracle(ctx, d, D1, D2, Dw1, Dw2, wl, wr, prec, tol):
    # an environment channel equal to the identity is applied as a copy of
    # the input (right side) or of the mixed slice into y (left side); the
    # oracle contracts it like any other channel
    W = sym_W(Dw1, Dw2, d, 60)
    L, R = rand(Dw1 * D1 * D1, 61), rand(Dw2 * D2 * D2, 62)
    if wl >= 0:
        L = with_identity_channel(L, Dw1, D1, wl)
    if wr >= 0:
        R = with_identity_channel(R, Dw2, D2, wr)
    x = rand(d * D1 * D2, 63)
    y_ref = O.heff_apply_ref(W, L, R, x, d, D1, D2)
    ctx.site_bind(W, L, R, d, D1, D2, prec)
    y = ctx.heff_apply(x)
    if prec == tn.TN_FP64:
        err = np.abs(y - y_ref).max() / np.abs(y_ref).max()
    else:
        err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    assert err < tol, err
    # a channel off the identity by more than the tolerance is contracted
    if wl >= 0:
        L2 = L.copy()
        L2[wl] += 1e-6  # entry (0, 0) of channel wl
        y2_ref = O.heff_apply_ref(W, L2, R, x, d, D1, D2)
        ctx.site_bind(W, L2, R, d, D1, D2, tn.TN_FP64)
        y2 = ctx.heff_apply(x)
        assert np.abs(y2 - y2_ref).max() / np.abs(y2_ref).max() < 1e-13


@pytest.mark.parametrize("d,D1,D2,Dw1,Dw2", HEFF_CASES)
def test_env_updates_match_oracle(ctx, d, D1, D2, Dw1, Dw2):
    W = sym_W(Dw1, Dw2, d, 30)
    L = rand(Dw1 * D1 * D1, 31)
    R = rand(Dw2 * D2 * D2, 32)
    M = rand(d * D1 * D2, 33)
    ctx.site_bind(W, L, R, d, D1, D2, tn.TN_FP64)
    Lo = ctx.env_update(0, M, Dw2 * D2 * D2)
    Lr = O.env_left_ref(W, L, M, d, D1, D2)
    assert np.abs(Lo - Lr).max() / np.abs(Lr).max() < 1e-13
    Ro = ctx.env_update(1, M, Dw1 * D1 * D1)
    Rr = O.env_right_ref(W, R, M, d, D1, D2)
    assert np.abs(Ro - Rr).max() / np.abs(Rr).max() < 1e-13


@pytest.mark.parametrize("d,D1,D2", [(2, 4, 6), (3, 9, 9), (2, 12, 10)])
def test_lanczos_lowest_matches_dense(ctx, d, D1, D2):
    W = sym_W(5, 5, d, 40)
    L, R = sym_env(5, D1, 41), sym_env(5, D2, 42)
    H = O.heff_dense(W, L, R, d, D1, D2)
    assert np.abs(H - H.T).max() < 1e-13  # the Hermiticity check of LapackEigenSolver.C:71-77
    w = np.linalg.eigvalsh(H)
    ctx.site_bind(W, L, R, d, D1, D2, tn.TN_FP64)
    ev, x, nmv, rc = ctx.eig_lowest(rand(d * D1 * D2, 43), tol=1e-11, maxit=400, nev=1)
    assert rc == 0
    assert abs(ev[0] - w[0]) < 1e-10 * max(1, abs(w[0]))
    assert abs(np.linalg.norm(x) - 1) < 1e-12
    assert np.linalg.norm(H @ x - ev[0] * x) < 1e-9


@pytest.mark.parametrize("m,n,rank", [(40, 24, 24), (24, 40, 24), (64, 64, 64), (96, 30, 7), (5, 3, 3),
                                      (1, 8, 1), (130, 100, 100)])
def test_svd_matches_numpy(ctx, m, n, rank):
    A = rand((m, rank), 50) @ rand((rank, n), 51)
    U, s, Vt = ctx.svd(A)
    k = min(m, n)
    s_np = np.linalg.svd(A, compute_uv=False)
    assert np.abs(s - s_np).max() < 1e-12 * s_np[0]
    assert np.abs(U @ np.diag(s) @ Vt - A).max() < 1e-12 * s_np[0]
    assert np.abs(U.T @ U - np.eye(k)).max() < 1e-12
    assert np.abs(Vt @ Vt.T - np.eye(k)).max() < 1e-12
    assert np.all(np.diff(s) <= 1e-15 * s_np[0])


@pytest.mark.parametrize("m,n,decay,rank", [(512, 512, 25.0, 512), (300, 160, 10.0, 160), (160, 300, 1e9, 160),
                                            (256, 256, 8.0, 200), (520, 200, 30.0, 200), (129, 700, 5.0, 129),
                                            (1024, 1024, 40.0, 1024), (2048, 384, 20.0, 300), (256, 128, 8.0, 100)])
def test_svd_graded_spectrum(ctx, m, n, decay, rank):
    """min(m, n) >= 128 takes the QR-preconditioned Jacobi.  Singular values
    exp(-j/decay) down to ~1e-28 (a two-site split's spectrum decays like this),
    exact zeros past the rank, tall and wide: the values must equal the
    constructed ones and both factors stay orthonormal."""
    rng = np.random.default_rng(60 + m + n)
    k = min(m, n)
    Q1, _ = np.linalg.qr(rng.standard_normal((m, k)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, k)))
    sig = np.exp(-np.arange(k) / decay)
    sig[rank:] = 0.0
    A = (Q1 * sig) @ Q2.T
    U, s, Vt = ctx.svd(A)
    assert np.abs(s - sig).max() < 1e-12
    assert np.abs(U @ np.diag(s) @ Vt - A).max() < 1e-12
    assert np.abs(U.T @ U - np.eye(k)).max() < 1e-12
    assert np.abs(Vt @ Vt.T - np.eye(k)).max() < 1e-12
    assert np.all(np.diff(s) <= 1e-15)


@pytest.mark.parametrize("m,n", [(300, 200), (256, 512)])
def test_svd_nearly_orthogonal_columns(ctx, m, n):
    """Columns (rows, when wide) already nearly orthogonal, the shape of a
    one-site gauge move's site tensor: the plain iteration runs on them.
    Pinned against numpy's singular values and by orthonormality."""
    rng = np.random.default_rng(70 + m)
    k = min(m, n)
    Q1, _ = np.linalg.qr(rng.standard_normal((max(m, n), k)))
    sig = np.exp(-np.arange(k) / 20.0)
    X = (Q1 * sig) @ (np.eye(k) + 1e-3 * rng.standard_normal((k, k)))
    A = X if m >= n else X.T
    U, s, Vt = ctx.svd(A)
    s_np = np.linalg.svd(A, compute_uv=False)
    assert np.abs(s - s_np).max() < 1e-12
    assert np.abs(U @ np.diag(s) @ Vt - A).max() < 1e-12
    assert np.abs(U.T @ U - np.eye(k)).max() < 1e-12
    assert np.abs(Vt @ Vt.T - np.eye(k)).max() < 1e-12


def _qr_positive(A):
    """Householder QR with R's diagonal made positive: the unique QR of a full
    column rank A, which every correct QR must reproduce."""
    q, r = np.linalg.qr(A)
    sg = np.sign(np.diag(r))
    sg[sg == 0] = 1.0
    return q * sg, (r.T * sg).T


@pytest.mark.parametrize("m,n", [(64, 64), (200, 96), (300, 200), (640, 520), (1100, 1024)])
def test_cholqr_matches_householder(ctx, m, n):
    # well conditioned: the QR with positive diag(R) is unique, so the device's
    # shifted CholeskyQR must agree with LAPACK's Householder QR
    rng = np.random.default_rng(m * 7 + n)
    A = rng.standard_normal((m, n))
    Q, R, ok, shifted = ctx.cholqr_test(A)
    assert ok and not shifted
    q, r = _qr_positive(A)
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-12
    assert np.abs(np.tril(R, -1)).max() == 0.0
    assert np.abs(Q @ R - A).max() < 1e-12 * np.abs(A).max() * np.sqrt(m)
    assert np.abs(R - r).max() < 1e-10 * np.abs(r).max()
    assert np.abs(Q - q).max() < 1e-9


@pytest.mark.parametrize("cond", [1e10, 1e14])
def test_cholqr_ill_conditioned_uses_shift(ctx, cond):
    # A = U diag(s) V^T with s spanning `cond`: plain Cholesky of A^T A breaks
    # down, the shifted first pass keeps Q orthonormal and Q R = A
    rng = np.random.default_rng(5)
    m, n = 384, 256
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -np.log10(cond), n)
    A = (U * s) @ V.T
    Q, R, ok, shifted = ctx.cholqr_test(A)
    assert ok and shifted
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-11
    assert np.abs(Q @ R - A).max() < 1e-13
    assert np.abs(np.tril(R, -1)).max() == 0.0


def test_cholqr_rank_deficient(ctx):
    # exactly repeated columns (Schmidt values that are exactly zero)
    rng = np.random.default_rng(9)
    B = rng.standard_normal((256, 80))
    A = np.concatenate([B, B[:, :48]], axis=1)
    Q, R, ok, shifted = ctx.cholqr_test(A)
    assert ok and shifted
    assert np.abs(Q.T @ Q - np.eye(128)).max() < 1e-11
    assert np.abs(Q @ R - A).max() < 1e-12


@pytest.mark.parametrize("cond", [1.0, 1e2, 1e3, 1e4, 1e5, 1e6, 1e12])
def test_cholqr_polar_finish(ctx, cond):
    # the gauge move's mode: any A = Q R with orthonormal Q (R need not be
    # triangular); well-conditioned inputs take the polar finish (its series
    # order grows with the first pass's loss of orthogonality, ~u cond^2:
    # none, first, second or third order), the others the further
    # CholeskyQR passes
    rng = np.random.default_rng(int(np.log10(cond)) + 11)
    m, n = 1100, 700
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (U * np.logspace(0, -np.log10(cond), n)) @ V.T
    Q, R, ok, _ = ctx.cholqr_test(A, triangular=False)
    assert ok
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-11
    assert np.abs(Q @ R - A).max() < 1e-12 * np.abs(A).max() * np.sqrt(n)


# ---- bulk shapes of the benched workload and the split-bfloat16 product ----

@pytest.mark.parametrize("tin,tol", [(3, 3e-7), (1, 3e-7)])
@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4])
def test_split_bf16_gemm_reaches_fp32_accuracy(ctx, tin, tol, tile):
    # three bfloat16 pieces per operand and the six leading cross products:
    # against the exact (FP64) product of the unrounded operands the error is
    # FP32-class, far below a plain bfloat16 product (~1e-3)
    M, N, K = 256, 384, 512
    A = rand((1, M, K), 31)
    B = rand((1, K, N), 32)
    ref = A @ B
    C = ctx.gemm_test(A, np.ascontiguousarray(B.transpose(0, 2, 1)), tin, 1, True, True, M, N, K, 1, tile)
    err = np.abs(C - ref).max() / np.abs(ref).max()
    assert err < tol * np.sqrt(K), (tin, tile, err)
    if tin == 3:
        Cb = ctx.gemm_test(A, np.ascontiguousarray(B.transpose(0, 2, 1)), 2, 1, True, True, M, N, K, 1, tile)
        assert np.abs(Cb - ref).max() / np.abs(ref).max() > 30 * err


@pytest.mark.parametrize("M,N,K,tile,ks", [
    (384, 384, 256, 11, 1),    # 96 x 128 tiles, 4 x 3 blocks
    (3072, 512, 128, 11, 1),   # 32 row blocks of 96: four full groups of 8
    (256, 288, 512, 12, 1),    # 128 x 96 tiles
    (1024, 384, 1024, 12, 2),  # 128 x 96 tiles with split-K into FP64
    (384, 512, 256, 13, 1),    # eight-wave 192 x 256 tiles (96 x 64 per wave)
    (3072, 1024, 64, 13, 1),   # 16 row blocks of 192: two full groups of 8
    (512, 384, 512, 14, 1),    # eight-wave 256 x 192 tiles (64 x 96 per wave)
    (1024, 384, 1024, 14, 4),  # 256 x 192 tiles with a 4-way split-K into FP64
    (384, 256, 256, 15, 1),    # eight waves, two LDS stages: 192 x 128
    (3072, 512, 128, 15, 1),   # 16 row blocks of 192: two full groups of 8
    (256, 384, 512, 16, 1),    # eight waves, two LDS stages: 128 x 192
    (1024, 384, 96, 16, 1),    # three K steps (an odd count: the stages alternate)
    (1024, 384, 1024, 16, 2),  # 128 x 192 with a 2-way split-K into FP64
])
def test_split_bf16_96_wide_tiles(ctx, M, N, K, tile, ks):
    # the 96-row / 96-column tiles of the split-bfloat16 product (three row
    # or column fragments per wave) and its eight-wave tiles; the other
    # element types refuse them
    A = rand((1, M, K), M + 7)
    B = rand((1, K, N), N + 8)
    ref = A @ B
    Bt = np.ascontiguousarray(B.transpose(0, 2, 1))
    C = ctx.gemm_test(A, Bt, 3, 0 if ks > 1 else 1, True, True, M, N, K, 1, tile, ks)
    err = np.abs(C - ref).max() / np.abs(ref).max()
    assert err < 3e-7 * np.sqrt(K), (M, N, K, tile, err)
    for tin in (1, 2):
        with pytest.raises(tn.TnError):
            ctx.gemm_test(A, Bt, tin, 1, True, True, M, N, K, 1, tile)


@pytest.mark.parametrize("tin", [1, 2, 3])
@pytest.mark.parametrize("M,N,K,tile", [
    (3072, 256, 256, 1),   # 24 row blocks of 128: three full groups of 8 (the bulk stage-1 row count)
    (1664, 256, 128, 1),   # 13 row blocks: one group of 8 and a partial group of 5
    (1664, 96, 64, 7),     # 52 blocks of 32 rows, block count not a multiple of 8 (no XCD remap)
    (2176, 640, 96, 3),    # 34 row blocks of 64 x 10 column blocks
])
def test_gemm_tile_order_many_row_blocks(ctx, tin, M, N, K, tile):
    # the XCD-aware, grouped tile order of gemm_k with more than 8 row blocks,
    # partial last groups and block counts that are not a multiple of 8
    A = rand((1, M, K), M + 1)
    B = rand((1, K, N), N + 2)
    ref = (ROUND[tin] if tin != 3 else (lambda x: x))(A) @ (ROUND[tin] if tin != 3 else (lambda x: x))(B)
    C = ctx.gemm_test(A, np.ascontiguousarray(B.transpose(0, 2, 1)), tin, 1, True, True, M, N, K, 1, tile)
    err = np.abs(C - ref).max() / np.abs(ref).max()
    assert err < (3e-7 * np.sqrt(K)), (tin, M, N, err)


@pytest.mark.parametrize("tin", [1, 2, 3])
@pytest.mark.parametrize("ks", [2, 4])
def test_split_k_into_fp64_output(ctx, tin, ks):
    # stage 3 of the benched apply: K split into ks chunks, partial products
    # in FP32, reduced into an FP64 output
    M, N, K = 1024, 384, 2048
    A = rand((1, M, K), 41)
    B = rand((1, K, N), 42)
    Ar = A if tin == 3 else ROUND[tin](A)
    Br = B if tin == 3 else ROUND[tin](B)
    ref = Ar @ Br
    C = ctx.gemm_test(A, np.ascontiguousarray(B.transpose(0, 2, 1)), tin, 0, True, True, M, N, K, 1, 0, ks)
    err = np.abs(C - ref).max() / np.abs(ref).max()
    assert err < 3e-7 * np.sqrt(K), (tin, ks, err)


@pytest.mark.parametrize("D", [256, 512])
@pytest.mark.parametrize("ident", [False, True])
@pytest.mark.parametrize("prec,tol", [(tn.TN_FP32, 5e-6), (tn.TN_BF16X3, 5e-6), (tn.TN_BF16, 2e-2)])
def test_heff_apply_bulk_shapes_match_oracle(ctx, D, ident, prec, tol):
    # the benched operator shapes (d=3, Dw=5, D1 = D2 = D): stage 1 is
    # (3D x 4D or 5D) x D, stage 3 D x 3D x (4D or 5D), with and without the
    # identity channels copied
    d = 3
    W = np.ascontiguousarray(O.xxz_bulk_W(d, 1.0, 1.0, 0.0, 0.1, 0))
    L, R = rand(5 * D * D, 51 + D), rand(5 * D * D, 52 + D)
    if ident:
        L = with_identity_channel(L, 5, D, 0)
        R = with_identity_channel(R, 5, D, 4)
    x = rand(d * D * D, 53)
    # the oracle's GEMM-ordered apply (OpenMP; pinned against the reference-
    # layout apply in tests/test_oracle.py) keeps D=512 within seconds
    y_ref = O.psi_int_to_ref(O.heff_apply_int(W, O.lenv_ref_to_int(L, 5, D), O.renv_ref_to_int(R, 5, D),
                                              O.psi_ref_to_int(x, d, D, D), d, D, D), d, D, D)
    ctx.site_bind(W, L, R, d, D, D, prec)
    y = ctx.heff_apply(x)
    err = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    assert err < tol, err
