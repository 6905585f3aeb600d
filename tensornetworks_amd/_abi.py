"""ctypes binding of include/tn_abi.h (libtn_hip.so)."""
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtn_hip.so")

TN_FP64, TN_FP32, TN_BF16, TN_BF16X3 = 0, 1, 2, 3
TN_OK, TN_WARN_NOT_CONVERGED = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)


class TnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tn error {code}: {msg}")
        self.code = code


class SweepOpts(ctypes.Structure):
    _fields_ = [
        ("eig_tol", ctypes.c_double),
        ("max_krylov", ctypes.c_int),
        ("max_restarts", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("env_precision", ctypes.c_int),
        ("refine", ctypes.c_int),
        ("refine_tol", ctypes.c_double),
        ("gemm_tile", ctypes.c_int),
        ("gauge", ctypes.c_int),
        ("refine_precision", ctypes.c_int),
        ("eig_method", ctypes.c_int),
        ("precond", ctypes.c_int),
        ("precond_delta", ctypes.c_double),
        ("skip_identity", ctypes.c_int),
        ("time_heff", ctypes.c_int),
        ("reorth", ctypes.c_int),
    ]

    @classmethod
    def default(cls, **kw):
        """tn_sweep_opts_default, then the given fields; an unknown field name
        raises (a ctypes structure would keep it as a plain attribute)."""
        o = cls()
        lib.tn_sweep_opts_default(ctypes.byref(o))
        names = {f for f, _ in cls._fields_}
        for k, v in kw.items():
            if k not in names:
                raise TypeError(f"tn_sweep_opts has no field {k!r}")
            setattr(o, k, v)
        return o


class SweepStats(ctypes.Structure):
    _fields_ = [
        ("n_matvec", ctypes.c_int64),
        ("n_matvec_fp64", ctypes.c_int64),
        ("n_env", ctypes.c_int64),
        ("n_svd", ctypes.c_int64),
        ("heff_flops", ctypes.c_double),
        ("heff_ms", ctypes.c_double),
        ("heff_gemm_ms", ctypes.c_double),
        ("heff_gemm_flops", ctypes.c_double),
        ("env_flops", ctypes.c_double),
        ("env_ms", ctypes.c_double),
        ("svd_ms", ctypes.c_double),
        ("lanczos_ms", ctypes.c_double),
        ("wall_ms", ctypes.c_double),
        ("energy", ctypes.c_double),
        ("n_qr", ctypes.c_int64),
        ("n_qr_fallback", ctypes.c_int64),
        ("n_qr_shifted", ctypes.c_int64),
        ("n_ident_left", ctypes.c_int64),
        ("n_ident_right", ctypes.c_int64),
        ("n_precond_fallback", ctypes.c_int64),
        ("n_full_reorth", ctypes.c_int64),
        ("n_solve_converged", ctypes.c_int64),
        ("n_solve_capped", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python tensornetworks_amd/build.py` "
            "(the HIP library is required: there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    i, d, v, i64, u64 = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64
    vp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "tn_version": (i, []),
        "tn_device_count": (i, []),
        "tn_ctx_create": (i, [i, vp]),
        "tn_ctx_destroy": (i, [v]),
        "tn_last_error": (ctypes.c_char_p, [v]),
        "tn_site_bind": (i, [v, i, i, i, i, i, _dp, _dp, _dp, i]),
        "tn_heff_apply": (i, [v, _dp, _dp, i64]),
        "tn_env_update": (i, [v, i, _dp, _dp]),
        "tn_eig_lowest": (i, [v, i64, i, d, i, _dp, _dp, _ip]),
        "tn_svd": (i, [v, i, i, _dp, _dp, _dp, _dp]),
        "tn_svcompress_rank": (i, [i, _dp, i, d, _dp, _dp]),
        "tn_gemm_test": (i, [v, i, i, i, i, i, i, i, i, _dp, _dp, _dp, i, i]),
        "tn_cholqr_test": (i, [v, i, i, _dp, ctypes.c_double, _dp, _dp, _ip, i]),
        "tn_ctx_set_autotune": (i, [v, i]),
        "tn_copy_bench": (i, [v, ctypes.c_int64, _dp]),
        "tn_ctx_tune_load": (i, [v, ctypes.c_char_p]),
        "tn_ctx_set_mlp": (i, [v, _fp, _fp]),
        "tn_mlp_forward": (i, [v, _fp, _fp, ctypes.c_int64, _dp, _dp]),
        "tn_mlp_loss_grad": (i, [v, _fp, _fp, ctypes.c_int64, _dp, _dp, _dp, _dp, _fp]),
        "tn_mlp_train": (i, [v, _fp, _fp, ctypes.c_int64, _dp, _dp, _dp, i, i, ctypes.c_double, u64, _dp]),
        "tn_chain_precond_samples": (i, [v, i, i, i, u64, ctypes.c_double, _dp, _dp, _dp, _ip]),
        "tn_batch_create": (i, [v, i, ctypes.POINTER(ctypes.c_void_p), i, vp]),
        "tn_batch_destroy": (i, [v]),
        "tn_batch_sweep": (i, [v, i, ctypes.POINTER(SweepOpts), ctypes.POINTER(SweepStats), _dp]),
        "tn_ctx_tune_save": (i, [v, ctypes.c_char_p]),
        "tn_ctx_tune_report": (i, [v, ctypes.c_char_p, i]),
        "tn_chain_heff_bench": (i, [v, i, i, i, i, i, _dp, _dp, _dp, _dp, _dp]),
        "tn_chain_create": (i, [v, i, i, _ip, _ip, _dp, i, i, vp]),
        "tn_chain_destroy": (i, [v]),
        "tn_chain_initialize": (i, [v, i, u64]),
        "tn_chain_get_dims": (i, [v, _ip, _ip]),
        "tn_chain_get_site": (i, [v, i, _dp]),
        "tn_chain_set_site": (i, [v, i, i, i, _dp]),
        "tn_sweep_opts_default": (None, [ctypes.POINTER(SweepOpts)]),
        "tn_chain_normalize": (i, [v, i]),
        "tn_chain_normalize_qr": (i, [v, i]),
        "tn_chain_load_caches": (i, [v, ctypes.POINTER(SweepOpts)]),
        "tn_chain_sweep": (i, [v, i, ctypes.POINTER(SweepOpts), ctypes.POINTER(SweepStats)]),
        "tn_chain_refine_site": (i, [v, i, i, ctypes.POINTER(SweepOpts), ctypes.POINTER(SweepStats)]),
        "tn_chain_increase_bond_dims": (i, [v, i]),
        "tn_chain_deltas": (i, [v, _dp, _dp]),
        "tn_chain_bond_sv": (i, [v, i, _dp, i]),
        "tn_chain_expectation": (i, [v, _ip, _ip, _dp, _dp]),
        "tn_chain_sweep2": (i, [v, i, i, d, ctypes.POINTER(SweepOpts), ctypes.POINTER(SweepStats), _dp]),
        "tn_heff_apply_device": (i, [v, v, v, i64]),
        "tn_heff_apply_device_noalloc": (i, [v, v, v, i64]),
        "tn_qr": (i, [v, i, i, i, _dp, _dp, _dp]),
        "tn_publish_test": (i, [v, i, _dp, _dp]),
        "tn_chain_norm_status": (i, [v, d, ctypes.c_char_p]),
        "tn_chain_save": (i, [v, ctypes.c_char_p, ctypes.c_char_p, i64, i]),
        "tn_chain_load": (i, [v, ctypes.c_char_p, i, vp, ctypes.c_char_p, i64, ctypes.POINTER(ctypes.c_int64)]),
        "tn_mpo_reduce": (i, [i, i, _ip, _ip, _dp, i, i, i, d, _ip, _ip, _dp, i64, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


MLP_NP = 4545  # parameters of the 4 -> 64 -> 64 -> 1 preconditioner network


def FP(a):
    """numpy float32 C-contiguous array -> float*"""
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_fp)


def P(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def IP(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def svcompress_rank(s, Dmax, eps):
    """SVCompressorImp::Compress rule -> (D, discarded, rescale)."""
    s = np.ascontiguousarray(s, dtype=np.float64)
    disc, resc = ctypes.c_double(), ctypes.c_double()
    D = lib.tn_svcompress_rank(len(s), P(s), Dmax, eps, ctypes.byref(disc), ctypes.byref(resc))
    if D < 0:
        raise TnError(D, "tn_svcompress_rank")
    return D, disc.value, resc.value


class Context:
    """One HIP stream + scratch; the owner of every device call."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib.tn_ctx_create(device, ctypes.byref(h))
        if rc != 0:
            raise TnError(rc, f"tn_ctx_create(device={device}) failed")
        self.h = h
        self._chains = weakref.WeakSet()

    def close(self):
        """Destroy the context; chains created on it are destroyed first."""
        if getattr(self, "h", None):
            for ch in list(getattr(self, "_chains", ())):
                ch.close()
            lib.tn_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_autotune(self, on=True):
        """Per-shape GEMM tile / k-split autotuning for this context."""
        self.check(lib.tn_ctx_set_autotune(self.h, int(bool(on))), "tn_ctx_set_autotune")

    def tune_load(self, path):
        """Merge a saved tuning table; returns the number of entries."""
        return self.check(lib.tn_ctx_tune_load(self.h, os.fsencode(path)), "tn_ctx_tune_load")

    def tune_save(self, path):
        return self.check(lib.tn_ctx_tune_save(self.h, os.fsencode(path)), "tn_ctx_tune_save")

    # -- learned Davidson preconditioner ---------------------------------------
    def set_mlp(self, params, norm):
        """Network used by SweepOpts(eig_method=1, precond=2 or 3) on this context."""
        p = np.ascontiguousarray(params, dtype=np.float32)
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        self.check(lib.tn_ctx_set_mlp(self.h, FP(p), FP(nm)), "tn_ctx_set_mlp")

    def mlp_forward(self, params, norm, X):
        X = np.ascontiguousarray(X, dtype=np.float64)
        g = np.zeros(X.shape[0])
        self.check(lib.tn_mlp_forward(self.h, FP(np.ascontiguousarray(params, dtype=np.float32)),
                                      FP(np.ascontiguousarray(norm, dtype=np.float32)), X.shape[0], P(X), P(g)),
                   "tn_mlp_forward")
        return g

    def mlp_loss_grad(self, params, norm, X, a, t):
        X = np.ascontiguousarray(X, dtype=np.float64)
        grad = np.zeros(MLP_NP, dtype=np.float32)
        loss = ctypes.c_double(0.0)
        self.check(lib.tn_mlp_loss_grad(self.h, FP(np.ascontiguousarray(params, dtype=np.float32)),
                                        FP(np.ascontiguousarray(norm, dtype=np.float32)), X.shape[0], P(X),
                                        P(np.ascontiguousarray(a, dtype=np.float64)),
                                        P(np.ascontiguousarray(t, dtype=np.float64)), ctypes.byref(loss), FP(grad)),
                   "tn_mlp_loss_grad")
        return loss.value, grad

    def mlp_train(self, params, norm, X, a, t, steps, batch=4096, lr=1e-3, seed=0):
        """Adam on the device; returns (new params, (loss before, loss after))."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        p = np.array(params, dtype=np.float32, copy=True)
        loss = np.zeros(2)
        self.check(lib.tn_mlp_train(self.h, FP(p), FP(np.ascontiguousarray(norm, dtype=np.float32)), X.shape[0],
                                    P(X), P(np.ascontiguousarray(a, dtype=np.float64)),
                                    P(np.ascontiguousarray(t, dtype=np.float64)), int(steps), int(batch), float(lr),
                                    int(seed), P(loss)), "tn_mlp_train")
        return p, (loss[0], loss[1])

    def copy_bench(self, n):
        """Duration (ms) of one plain device copy of n doubles."""
        ms = ctypes.c_double(0.0)
        self.check(lib.tn_copy_bench(self.h, int(n), ctypes.byref(ms)), "tn_copy_bench")
        return ms.value

    def tune_report(self):
        buf = ctypes.create_string_buffer(1 << 16)
        lib.tn_ctx_tune_report(self.h, buf, len(buf))
        return buf.value.decode()

    def check(self, rc, what):
        if rc < 0:
            msg = lib.tn_last_error(self.h)
            raise TnError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    # -- reference seam -------------------------------------------------------
    def site_bind(self, W, L, R, d, D1, D2, precision=TN_FP64):
        W = np.ascontiguousarray(W, dtype=np.float64)
        Dw1, Dw2 = W.shape[0], W.shape[1]
        self._bound = (np.ascontiguousarray(L, dtype=np.float64), np.ascontiguousarray(R, dtype=np.float64), W)
        self.check(lib.tn_site_bind(self.h, d, D1, D2, Dw1, Dw2, P(W), P(self._bound[0]), P(self._bound[1]), precision),
                   "tn_site_bind")

    def heff_apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self.check(lib.tn_heff_apply(self.h, P(x), P(y), x.size), "tn_heff_apply")
        return y

    def env_update(self, direction, M, out_size):
        M = np.ascontiguousarray(M, dtype=np.float64)
        F = np.zeros(out_size)
        self.check(lib.tn_env_update(self.h, direction, P(M), P(F)), "tn_env_update")
        return F

    def eig_lowest(self, x0, tol=1e-12, maxit=200, nev=2):
        """Lowest nev (1 or 2) eigenpairs of the bound operator from start
        vector x0: returns (evals, X, n_matvec, rc) with X[:, k] the
        eigenvector of evals[k] (X is 1-D, the first vector, when nev == 1)."""
        x0 = np.ascontiguousarray(x0, dtype=np.float64).ravel()
        N = x0.size
        x = np.zeros(nev * N)
        x[:N] = x0
        ev = np.zeros(2)
        nmv = ctypes.c_int()
        rc = self.check(lib.tn_eig_lowest(self.h, N, nev, tol, maxit, P(ev), P(x), ctypes.byref(nmv)),
                        "tn_eig_lowest")
        X = x[:N].copy() if nev == 1 else x.reshape(nev, N).T.copy()
        return ev[:nev], X, nmv.value, rc

    def heff_apply_device(self, x_ptr, y_ptr, n):
        """y = Heff x with x, y device pointers (reference layout, n doubles)."""
        self.check(lib.tn_heff_apply_device(self.h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), int(n)),
                   "tn_heff_apply_device")

    def heff_apply_device_noalloc(self, x_ptr, y_ptr, n):
        """tn_heff_apply_device through the fixed work space (allocates nothing)."""
        self.check(lib.tn_heff_apply_device_noalloc(self.h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(y_ptr), int(n)),
                   "tn_heff_apply_device_noalloc")

    def qr(self, A, mode=0):
        """Thin QR (mode 0, m >= n: A = Q R) or LQ (mode 1, m <= n: A = L Q) on
        the device; returns (Q, R or L, rc)."""
        A = np.asarray(A, dtype=np.float64)
        m, n = A.shape
        Af = np.ascontiguousarray(A.T).ravel()
        k = n if mode == 0 else m
        Q = np.zeros(m * n)
        R = np.zeros(k * k)
        rc = self.check(lib.tn_qr(self.h, int(mode), m, n, P(Af), P(Q), P(R)), "tn_qr")
        return Q.reshape(n, m).T.copy(), R.reshape(k, k).T.copy(), rc

    def publish_test(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self.check(lib.tn_publish_test(self.h, x.size, P(x), P(y)), "tn_publish_test")
        return y

    def svd(self, A):
        A = np.asarray(A, dtype=np.float64)
        m, n = A.shape
        k = min(m, n)
        Af = np.ascontiguousarray(A.T).ravel()  # column-major
        U, s, Vt = np.zeros(m * k), np.zeros(k), np.zeros(k * n)
        self.check(lib.tn_svd(self.h, m, n, P(Af), P(U), P(s), P(Vt)), "tn_svd")
        return U.reshape(k, m).T.copy(), s, Vt.reshape(n, k).T.copy()

    def gemm_test(self, A, B, tin, tout, a_kmajor, b_kmajor, M, N, K, batch=1, tile=0, ksplit=1):
        A = np.ascontiguousarray(A, dtype=np.float64).ravel()
        B = np.ascontiguousarray(B, dtype=np.float64).ravel()
        C = np.zeros(batch * M * N)
        self.check(lib.tn_gemm_test(self.h, tin, tout, int(a_kmajor), int(b_kmajor), M, N, K, batch, P(A), P(B),
                                    P(C), tile, ksplit), "tn_gemm_test")
        return C.reshape(batch, M, N)

    def cholqr_test(self, A, tol=1e-11, triangular=True):
        """A = Q R on the device (shifted CholeskyQR); returns (Q, R, ok, shifted).
        triangular=False allows the polar finish used by the gauge move."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        m, n = A.shape
        Q = np.zeros((m, n))
        R = np.zeros((n, n))
        sh = ctypes.c_int(0)
        rc = lib.tn_cholqr_test(self.h, m, n, P(A.ravel()), tol, P(Q.ravel()), P(R.ravel()), ctypes.byref(sh),
                                int(bool(triangular)))
        if rc < 0:
            self.check(rc, "tn_cholqr_test")
        return Q, R, rc == 0, sh.value


def _mpo_arrays(Ws):
    dw1 = np.array([W.shape[0] for W in Ws], dtype=np.int32)
    dw2 = np.array([W.shape[1] for W in Ws], dtype=np.int32)
    wcat = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(W, dtype=np.float64).ravel() for W in Ws]))
    return dw1, dw2, wcat


def read_checkpoint_mpo(path):
    """The MPO stored in a chain checkpoint (list of W[i], (Dw1, Dw2, d, d))."""
    if not os.path.isfile(path):
        raise TnError(-2, f"{path}: no such checkpoint file")
    with open(path, "rb") as f:
        head = f.read(8)
        if head != b"TNCHKPT1":
            raise TnError(-2, f"{path}: not a chain checkpoint")
        version, L, d, _dcap, _flags = np.frombuffer(f.read(20), dtype="<i4")
        if version != 1:
            raise TnError(-2, f"{path}: unknown checkpoint version {version}")
        dw = np.frombuffer(f.read(8 * L), dtype="<i4").reshape(2, L)
        Ws = []
        for i in range(L):
            n = int(dw[0, i] * dw[1, i] * d * d)
            Ws.append(np.frombuffer(f.read(8 * n), dtype="<f8").reshape(dw[0, i], dw[1, i], d, d).copy())
    return Ws


def mpo_reduce(Ws, form=0, ops=3, Dmax=0, eps=1e-13):
    """MPO::CanonicalForm (ops & 1) and MPO::Compress(Parker, Dmax, eps)
    (ops & 2) of a regular MPO (form 0 upper, 1 lower) through tn_mpo_reduce;
    returns (list of W[i], truncation error)."""
    dw1, dw2, wcat = _mpo_arrays(Ws)
    L = len(Ws)
    d = Ws[0].shape[2]
    o1 = np.zeros(L, dtype=np.int32)
    o2 = np.zeros(L, dtype=np.int32)
    te = ctypes.c_double(0.0)
    need = lib.tn_mpo_reduce(L, d, IP(dw1), IP(dw2), P(wcat), form, ops, Dmax, eps, IP(o1), IP(o2), None, 0,
                             ctypes.byref(te))
    if need < 0:
        raise TnError(need, "tn_mpo_reduce: bad MPO (not in regular form?)")
    out = np.zeros(max(need, 1))
    got = lib.tn_mpo_reduce(L, d, IP(dw1), IP(dw2), P(wcat), form, ops, Dmax, eps, IP(o1), IP(o2), P(out), out.size,
                            ctypes.byref(te))
    if got != need:
        raise TnError(-3, "tn_mpo_reduce: size changed between calls")
    Wo, off = [], 0
    for i in range(L):
        n = int(o1[i] * o2[i] * d * d)
        Wo.append(out[off:off + n].reshape(o1[i], o2[i], d, d).copy())
        off += n
    return Wo, te.value


class Chain:
    """Device-resident MPS of L sites with its MPO (list of W[i] arrays
    shaped (Dw1, Dw2, d, d)); sites start as in MPSImp::InitSitesAndBonds."""

    def __init__(self, ctx, Ws, D, Dcap=None):
        self.ctx = ctx
        self.Ws = Ws
        self.L = len(Ws)
        self.d = Ws[0].shape[2]
        self._mpo = _mpo_arrays(Ws)
        h = ctypes.c_void_p()
        dw1, dw2, wcat = self._mpo
        ctx.check(lib.tn_chain_create(ctx.h, self.L, self.d, IP(dw1), IP(dw2), P(wcat), D, Dcap or D, ctypes.byref(h)),
                  "tn_chain_create")
        self.h = h
        ctx._chains.add(self)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):
                lib.tn_chain_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def initialize(self, state=1, seed=20261202):
        self.ctx.check(lib.tn_chain_initialize(self.h, state, seed), "tn_chain_initialize")

    def dims(self):
        D1 = np.zeros(self.L, dtype=np.int32)
        D2 = np.zeros(self.L, dtype=np.int32)
        lib.tn_chain_get_dims(self.h, IP(D1), IP(D2))
        return D1.tolist(), D2.tolist()

    def site(self, i):
        D1, D2 = self.dims()
        M = np.zeros(self.d * D1[i] * D2[i])
        self.ctx.check(lib.tn_chain_get_site(self.h, i, P(M)), "tn_chain_get_site")
        return M

    def set_site(self, i, D1, D2, M):
        M = np.ascontiguousarray(M, dtype=np.float64)
        self.ctx.check(lib.tn_chain_set_site(self.h, i, D1, D2, P(M)), "tn_chain_set_site")

    def normalize(self, direction):
        self.ctx.check(lib.tn_chain_normalize(self.h, direction), "tn_chain_normalize")

    def normalize_qr(self, direction):
        self.ctx.check(lib.tn_chain_normalize_qr(self.h, direction), "tn_chain_normalize_qr")

    def load_caches(self, opts=None):
        opts = opts or SweepOpts.default()
        self.ctx.check(lib.tn_chain_load_caches(self.h, ctypes.byref(opts)), "tn_chain_load_caches")

    def sweep(self, direction, opts=None, stats=None):
        opts = opts or SweepOpts.default()
        st = stats if stats is not None else SweepStats()
        self.ctx.check(lib.tn_chain_sweep(self.h, direction, ctypes.byref(opts), ctypes.byref(st)), "tn_chain_sweep")
        return st

    def sweep2(self, direction, Dmax, eps_sv=0.0, opts=None, stats=None):
        opts = opts or SweepOpts.default()
        st = stats if stats is not None else SweepStats()
        tr = ctypes.c_double()
        self.ctx.check(lib.tn_chain_sweep2(self.h, direction, Dmax, eps_sv, ctypes.byref(opts), ctypes.byref(st),
                                           ctypes.byref(tr)), "tn_chain_sweep2")
        return st, tr.value

    def refine_site(self, i, direction, opts=None, stats=None):
        opts = opts or SweepOpts.default()
        st = stats if stats is not None else SweepStats()
        self.ctx.check(lib.tn_chain_refine_site(self.h, i, direction, ctypes.byref(opts), ctypes.byref(st)),
                       "tn_chain_refine_site")
        return st

    def increase_bond_dims(self, D):
        self.ctx.check(lib.tn_chain_increase_bond_dims(self.h, D), "tn_chain_increase_bond_dims")

    def deltas(self):
        de, dl = ctypes.c_double(), ctypes.c_double()
        lib.tn_chain_deltas(self.h, ctypes.byref(de), ctypes.byref(dl))
        return de.value, dl.value

    def bond_sv(self, b):
        buf = np.zeros(65536)
        n = lib.tn_chain_bond_sv(self.h, b, P(buf), buf.size)
        if n < 0:
            raise TnError(n, "tn_chain_bond_sv")
        return buf[:n].copy()

    def precond_samples(self, site, iters=4, nsamp=2048, seed=0, delta=1e-2):
        """Training samples (X, a, t) of the learned preconditioner at a site."""
        n = iters * nsamp
        X = np.zeros((n, 4))
        a = np.zeros(n)
        t = np.zeros(n)
        nout = ctypes.c_int(0)
        self.ctx.check(lib.tn_chain_precond_samples(self.h, site, iters, nsamp, int(seed), float(delta), P(X), P(a),
                                                    P(t), ctypes.byref(nout)), "tn_chain_precond_samples")
        k = nout.value
        return X[:k], a[:k], t[:k]

    def heff_bench(self, site, precision, reps=10, skip_identity=True, env_precision=None):
        """Average device ms of (GEMM 1, MPO stage, GEMM 2) and flops per GEMM
        launch, for the operator a sweep builds (identity environment channels
        applied as copies when skip_identity; env_precision defaults to the
        operator precision, FP32 for BF16)."""
        if env_precision is None:
            env_precision = TN_FP32 if precision == TN_BF16 else precision
        vals = [ctypes.c_double() for _ in range(5)]
        self.ctx.check(lib.tn_chain_heff_bench(self.h, site, precision, reps, int(bool(skip_identity)),
                                               env_precision, *[ctypes.byref(x) for x in vals]),
                       "tn_chain_heff_bench")
        return [x.value for x in vals]

    def norm_status(self, eps=1e-12):
        """MPS::GetNormStatus: one of 'A', 'B', 'I', 'M' per site."""
        buf = ctypes.create_string_buffer(self.L + 1)
        self.ctx.check(lib.tn_chain_norm_status(self.h, float(eps), buf), "tn_chain_norm_status")
        return buf.value.decode()

    def save(self, path, meta=b"", environments=False):
        """Checkpoint (tn_chain_save): MPO, sites, bond values, per-site
        energies, optionally the FP64 environments, and `meta` bytes."""
        meta = bytes(meta)
        self.ctx.check(lib.tn_chain_save(self.h, os.fsencode(path), meta, len(meta), 1 if environments else 0),
                       "tn_chain_save")

    @classmethod
    def load(cls, ctx, path, Dcap=0):
        """Chain from a checkpoint; returns (chain, meta bytes)."""
        Ws = read_checkpoint_mpo(path)
        h = ctypes.c_void_p()
        nb = ctypes.c_int64(0)
        buf = ctypes.create_string_buffer(1 << 20)
        ctx.check(lib.tn_chain_load(ctx.h, os.fsencode(path), int(Dcap), ctypes.byref(h), buf, len(buf),
                                    ctypes.byref(nb)), "tn_chain_load")
        if nb.value > len(buf):
            lib.tn_chain_destroy(h)
            raise TnError(-2, f"{path}: checkpoint meta data larger than {len(buf)} bytes")
        ch = cls.__new__(cls)
        ch.ctx = ctx
        ch.Ws = Ws
        ch.L = len(Ws)
        ch.d = Ws[0].shape[2]
        ch._mpo = _mpo_arrays(Ws)
        ch.h = h
        ctx._chains.add(ch)
        return ch, buf.raw[:nb.value]

    def expectation(self, Ws=None):
        dw1, dw2, wcat = self._mpo if Ws is None else _mpo_arrays(Ws)
        v = ctypes.c_double()
        self.ctx.check(lib.tn_chain_expectation(self.h, IP(dw1), IP(dw2), P(wcat), ctypes.byref(v)),
                       "tn_chain_expectation")
        return v.value


class Batch:
    """Many chains of one context swept in lockstep (tn_batch_*): every stage
    of a round is one grouped launch over the chains.  FP32 contractions, QR
    gauge; the chains must have their caches loaded and outlive the batch."""

    def __init__(self, ctx, chains, max_krylov=32):
        self.ctx = ctx
        self.chains = list(chains)
        arr = (ctypes.c_void_p * len(self.chains))(*[c.h.value for c in self.chains])
        h = ctypes.c_void_p()
        ctx.check(lib.tn_batch_create(ctx.h, len(self.chains), arr, int(max_krylov), ctypes.byref(h)),
                  "tn_batch_create")
        self.h = h
        ctx._chains.add(self)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):
                lib.tn_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def sweep(self, direction, opts, stats=None):
        """One pass of every chain; returns each chain's last Ritz value."""
        e = np.zeros(len(self.chains))
        self.ctx.check(lib.tn_batch_sweep(self.h, direction, ctypes.byref(opts),
                                          ctypes.byref(stats) if stats is not None else None, P(e)),
                       "tn_batch_sweep")
        return e
