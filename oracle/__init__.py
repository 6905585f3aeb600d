"""ctypes view of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this package, and only as the checker / timed baseline.
The product library never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtn_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with make (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    i, d, l, v, u64 = ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_void_p, ctypes.c_uint64
    sig = {
        "orc_version": (i, []),
        "orc_ed_ground_mpo": (d, [i, i, _ip, _ip, _dp, d, i]),
        "orc_spin_ops": (None, [i, _dp, _dp, _dp]),
        "orc_xxz_bulk_W": (None, [i, d, d, d, d, i, _dp]),
        "orc_site_W": (None, [i, i, _dp, i, i, _dp, _ip, _ip]),
        "orc_mpo_product": (None, [i, i, i, _dp, i, i, _dp, _dp]),
        "orc_heff_dense": (None, [i, i, i, i, i, _dp, _dp, _dp, _dp]),
        "orc_heff_apply_ref": (None, [i, i, i, i, i, _dp, _dp, _dp, _dp, _dp]),
        "orc_env_left_ref": (None, [i, i, i, i, i, _dp, _dp, _dp, _dp]),
        "orc_env_right_ref": (None, [i, i, i, i, i, _dp, _dp, _dp, _dp]),
        "orc_psi_ref_to_int": (None, [i, i, i, _dp, _dp]),
        "orc_psi_int_to_ref": (None, [i, i, i, _dp, _dp]),
        "orc_lenv_ref_to_int": (None, [i, i, _dp, _dp]),
        "orc_lenv_int_to_ref": (None, [i, i, _dp, _dp]),
        "orc_renv_ref_to_int": (None, [i, i, _dp, _dp]),
        "orc_renv_int_to_ref": (None, [i, i, _dp, _dp]),
        "orc_gemm": (None, [i, i, i, _dp, l, l, _dp, l, l, d, _dp, l]),
        "orc_heff_work_size": (l, [i, i, i, i, i]),
        "orc_heff_apply_int": (None, [i, i, i, i, i, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_env_left_int": (None, [i, i, i, i, i, _dp, _dp, _dp, _dp, _dp]),
        "orc_env_right_int": (None, [i, i, i, i, i, _dp, _dp, _dp, _dp, _dp]),
        "orc_syev": (i, [i, _dp, _dp, _dp]),
        "orc_svd": (i, [i, i, _dp, _dp, _dp, _dp]),
        "orc_svcompress": (i, [i, _dp, i, d, _dp, _dp]),
        "orc_ed_ground": (d, [i, i, d, d, d, d, d, i]),
        "orc_mps_create": (v, [i, i, d, d, d, d, i, i, i, u64]),
        "orc_mps_destroy": (None, [v]),
        "orc_mps_normalize": (None, [v, i]),
        "orc_mps_load_caches": (None, [v]),
        "orc_mps_find_gs_line": (d, [v, i, i, i, d, d, d, i, _ip]),
        "orc_mps_sweep": (None, [v, i, d, i]),
        "orc_mps_sweep2": (d, [v, i, i, d, d]),
        "orc_mps_expectation": (d, [v, i]),
        "orc_mps_get_dims": (i, [v, _ip, _ip]),
        "orc_mps_get_site": (i, [v, i, _dp]),
        "orc_mps_set_site": (i, [v, i, i, i, _dp]),
        "orc_mps_max_delta_e": (d, [v]),
        "orc_mps_max_delta_l": (d, [v]),
        "orc_mps_total_matvecs": (i, [v]),
        "orc_mps_last_energy": (d, [v]),
        "orc_mps_norm_status": (None, [v, ctypes.c_char_p]),
        "orc_mps_get_bond_sv": (i, [v, i, _dp]),
        "orc_rng_uniform": (d, [ctypes.POINTER(u64)]),
        "orc_mlp_features": (None, [l, _dp, d, _dp, _dp, d, _dp]),
        "orc_mlp_forward": (None, [_dp, _dp, l, _dp, _dp]),
        "orc_mlp_loss_grad": (d, [_dp, _dp, l, _dp, _dp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def P(a):
    """numpy float64 C-contiguous array -> double*"""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def spin_ops(d):
    sp, sm, sz = (np.zeros((d, d)) for _ in range(3))
    lib.orc_spin_ops(d, P(sp), P(sm), P(sz))
    return sp, sm, sz


def xxz_bulk_W(d, Jxy=1.0, Jz=1.0, hz=0.0, Dz=0.0, form=0):
    W = np.zeros((5, 5, d, d))
    lib.orc_xxz_bulk_W(d, Jxy, Jz, hz, Dz, form, P(W))
    return W


def site_W(d, Wbulk, form, pos):
    W = np.zeros(25 * d * d)
    a, b = ctypes.c_int(), ctypes.c_int()
    lib.orc_site_W(d, 5, P(np.ascontiguousarray(Wbulk)), form, pos, P(W), ctypes.byref(a), ctypes.byref(b))
    return W[: a.value * b.value * d * d].reshape(a.value, b.value, d, d).copy()


def heff_dense(W, L, R, d, D1, D2):
    Dw1, Dw2 = W.shape[0], W.shape[1]
    N = d * D1 * D2
    H = np.zeros((N, N))
    lib.orc_heff_dense(d, D1, D2, Dw1, Dw2, P(np.ascontiguousarray(W)), P(L), P(R), P(H))
    return H


def heff_apply_ref(W, L, R, x, d, D1, D2):
    y = np.zeros_like(x)
    lib.orc_heff_apply_ref(d, D1, D2, W.shape[0], W.shape[1], P(np.ascontiguousarray(W)), P(L), P(R), P(x), P(y))
    return y


def heff_apply_int(W, L, R, x, d, D1, D2):
    Dw1, Dw2 = W.shape[0], W.shape[1]
    work = np.zeros(lib.orc_heff_work_size(d, D1, D2, Dw1, Dw2))
    y = np.zeros(D1 * d * D2)
    lib.orc_heff_apply_int(d, D1, D2, Dw1, Dw2, P(np.ascontiguousarray(W)), P(L), P(R), P(x), P(y), P(work))
    return y


def env_left_ref(W, Fin, M, d, D1, D2):
    Dw1, Dw2 = W.shape[0], W.shape[1]
    out = np.zeros(Dw2 * D2 * D2)
    lib.orc_env_left_ref(d, D1, D2, Dw1, Dw2, P(np.ascontiguousarray(W)), P(Fin), P(M), P(out))
    return out


def env_right_ref(W, Fin, M, d, D1, D2):
    Dw1, Dw2 = W.shape[0], W.shape[1]
    out = np.zeros(Dw1 * D1 * D1)
    lib.orc_env_right_ref(d, D1, D2, Dw1, Dw2, P(np.ascontiguousarray(W)), P(Fin), P(M), P(out))
    return out


def env_left_int(W, Lin, A, d, D1, D2):
    Dw1, Dw2 = W.shape[0], W.shape[1]
    work = np.zeros(lib.orc_heff_work_size(d, D1, D2, Dw1, Dw2))
    out = np.zeros(D2 * Dw2 * D2)
    lib.orc_env_left_int(d, D1, D2, Dw1, Dw2, P(np.ascontiguousarray(W)), P(Lin), P(A), P(out), P(work))
    return out


def env_right_int(W, Rin, A, d, D1, D2):
    Dw1, Dw2 = W.shape[0], W.shape[1]
    work = np.zeros(lib.orc_heff_work_size(d, D1, D2, Dw1, Dw2))
    out = np.zeros(Dw1 * D1 * D1)
    lib.orc_env_right_int(d, D1, D2, Dw1, Dw2, P(np.ascontiguousarray(W)), P(Rin), P(A), P(out), P(work))
    return out


def psi_ref_to_int(x, d, D1, D2):
    y = np.zeros_like(x)
    lib.orc_psi_ref_to_int(d, D1, D2, P(x), P(y))
    return y


def psi_int_to_ref(x, d, D1, D2):
    y = np.zeros_like(x)
    lib.orc_psi_int_to_ref(d, D1, D2, P(x), P(y))
    return y


def lenv_ref_to_int(F, Dw, D):
    y = np.zeros_like(F)
    lib.orc_lenv_ref_to_int(Dw, D, P(F), P(y))
    return y


def lenv_int_to_ref(F, Dw, D):
    y = np.zeros_like(F)
    lib.orc_lenv_int_to_ref(Dw, D, P(F), P(y))
    return y


def renv_ref_to_int(F, Dw, D):
    y = np.zeros_like(F)
    lib.orc_renv_ref_to_int(Dw, D, P(F), P(y))
    return y


def renv_int_to_ref(F, Dw, D):
    y = np.zeros_like(F)
    lib.orc_renv_int_to_ref(Dw, D, P(F), P(y))
    return y


def syev(A):
    n = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float64)
    w = np.zeros(n)
    Z = np.zeros(n * n)
    lib.orc_syev(n, P(A), P(w), P(Z))
    return w, Z.reshape(n, n).T.copy()  # columns are eigenvectors


def svd(A):
    """A (m x n) numpy -> U (m x k), s (k), Vt (k x n)"""
    m, n = A.shape
    k = min(m, n)
    Af = np.asfortranarray(A, dtype=np.float64).ravel(order="F").copy()
    U = np.zeros(m * k)
    s = np.zeros(k)
    Vt = np.zeros(k * n)
    lib.orc_svd(m, n, P(Af), P(U), P(s), P(Vt))
    return U.reshape(k, m).T.copy(), s, Vt.reshape(n, k).T.copy()


def svcompress(s, Dmax, eps):
    s = np.ascontiguousarray(s, dtype=np.float64)
    disc, resc = ctypes.c_double(), ctypes.c_double()
    D = lib.orc_svcompress(len(s), P(s), Dmax, eps, ctypes.byref(disc), ctypes.byref(resc))
    return D, disc.value, resc.value


def ed_ground_mpo(Ws, tol=1e-12, maxit=300):
    """Ground energy of an MPO (list of W[i], (Dw1, Dw2, d, d)) by Lanczos on the d^L space."""
    dw1 = np.array([W.shape[0] for W in Ws], dtype=np.int32)
    dw2 = np.array([W.shape[1] for W in Ws], dtype=np.int32)
    wcat = np.ascontiguousarray(np.concatenate([np.ascontiguousarray(W, dtype=np.float64).ravel() for W in Ws]))
    return lib.orc_ed_ground_mpo(len(Ws), Ws[0].shape[2], dw1.ctypes.data_as(_ip), dw2.ctypes.data_as(_ip), P(wcat),
                                  tol, maxit)


def ed_ground(L, d, Jxy=1.0, Jz=1.0, hz=0.0, Dz=0.0, tol=1e-12, maxit=300):
    return lib.orc_ed_ground(L, d, Jxy, Jz, hz, Dz, tol, maxit)


class OracleMPS:
    """CPU restatement of MPSImp (one-site sweep, MPSImp_Variational.C:16-173)."""

    def __init__(self, L, d, D, Jxy=1.0, Jz=1.0, hz=0.0, Dz=0.0, form=0, init=1, seed=20261202):
        self.L, self.d = L, d
        self.h = lib.orc_mps_create(L, d, Jxy, Jz, hz, Dz, form, D, init, seed)

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_mps_destroy(self.h)
            self.h = None

    def normalize(self, direction):
        lib.orc_mps_normalize(self.h, direction)

    def load_caches(self):
        lib.orc_mps_load_caches(self.h)

    def find_ground_state(self, lines, mode=1):
        """lines: list of (maxiter, Dmax, DeltaD, epsE, epsL, epsEig)."""
        self.normalize(1)
        self.load_caches()
        de2 = 0.0
        ns = ctypes.c_int()
        for (maxiter, Dmax, dD, epsE, epsL, epsEig) in lines:
            de2 = lib.orc_mps_find_gs_line(self.h, maxiter, Dmax, dD, epsE, epsL, epsEig, mode, ctypes.byref(ns))
        self.nsweeps = ns.value
        return de2

    def sweep(self, direction, eps_eig=1e-12, mode=1):
        lib.orc_mps_sweep(self.h, direction, eps_eig, mode)

    def sweep2(self, direction, Dmax, eps_sv=0.0, eps_eig=1e-10):
        return lib.orc_mps_sweep2(self.h, direction, Dmax, eps_sv, eps_eig)

    def expectation(self, which=0):
        return lib.orc_mps_expectation(self.h, which)

    def dims(self):
        D1 = (ctypes.c_int * self.L)()
        D2 = (ctypes.c_int * self.L)()
        lib.orc_mps_get_dims(self.h, D1, D2)
        return list(D1), list(D2)

    def site(self, i):
        D1, D2 = self.dims()
        M = np.zeros(self.d * D1[i] * D2[i])
        lib.orc_mps_get_site(self.h, i, P(M))
        return M

    def set_site(self, i, D1, D2, M):
        lib.orc_mps_set_site(self.h, i, D1, D2, P(np.ascontiguousarray(M, dtype=np.float64)))

    def norm_status(self):
        buf = ctypes.create_string_buffer(self.L + 1)
        lib.orc_mps_norm_status(self.h, buf)
        return buf.value.decode()

    def max_delta_e(self):
        return lib.orc_mps_max_delta_e(self.h)

    def max_delta_l(self):
        return lib.orc_mps_max_delta_l(self.h)

    def matvecs(self):
        return lib.orc_mps_total_matvecs(self.h)

    def last_energy(self):
        return lib.orc_mps_last_energy(self.h)

    def bond_sv(self, b):
        sv = np.zeros(4096)
        n = lib.orc_mps_get_bond_sv(self.h, b, P(sv))
        return sv[:n].copy()


MLP_H = 64
MLP_NP = MLP_H * 4 + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1


def mlp_features(diag, theta, lam_l, lam_r, delta=1e-2):
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    X = np.zeros((diag.size, 4))
    lib.orc_mlp_features(diag.size, P(diag), float(theta), P(np.ascontiguousarray(lam_l, dtype=np.float64)),
                         P(np.ascontiguousarray(lam_r, dtype=np.float64)), float(delta), P(X))
    return X


def mlp_forward(params, norm, X):
    X = np.ascontiguousarray(X, dtype=np.float64)
    g = np.zeros(X.shape[0])
    lib.orc_mlp_forward(P(np.ascontiguousarray(params, dtype=np.float64)),
                        P(np.ascontiguousarray(norm, dtype=np.float64)), X.shape[0], P(X), P(g))
    return g


def mlp_loss_grad(params, norm, X, a, t):
    X = np.ascontiguousarray(X, dtype=np.float64)
    grad = np.zeros(MLP_NP)
    loss = lib.orc_mlp_loss_grad(P(np.ascontiguousarray(params, dtype=np.float64)),
                                 P(np.ascontiguousarray(norm, dtype=np.float64)), X.shape[0], P(X),
                                 P(np.ascontiguousarray(a, dtype=np.float64)),
                                 P(np.ascontiguousarray(t, dtype=np.float64)), P(grad))
    return loss, grad
