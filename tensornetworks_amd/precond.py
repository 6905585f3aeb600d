"""Learned Davidson preconditioner: data, training and storage.

The network (4 -> 64 -> 64 -> 1, tanh, FP32) outputs a factor g on the
diagonal Davidson correction t = g r / (theta - diag); its forward and
backward passes and the Adam steps run in the library's own kernels
(tn_mlp_train), the samples come from the method itself
(tn_chain_precond_samples: exact local corrections at small bond dimension).
With w3 = 0 and b3 = 1 the network is the plain diagonal preconditioner,
which is where training starts.
"""
import numpy as np

from ._abi import MLP_NP, SweepOpts, TN_FP64, Chain
from .models import xxz_mpo, batched_workload

H = 64
OFF_B1 = H * 4
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H


def init_params(seed=0):
    """Xavier-scaled hidden layers, output layer zero with bias 1 (g = 1)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_NP, dtype=np.float32)
    p[:OFF_B1] = rng.normal(0.0, np.sqrt(1.0 / 4.0), H * 4)
    p[OFF_W2:OFF_B2] = rng.normal(0.0, np.sqrt(1.0 / H), H * H)
    p[OFF_B3] = 1.0
    return p


def fit_norm(X):
    """Standardisation of the raw features: [mu(4), 1/sd(4)]."""
    mu = X.mean(axis=0)
    sd = X.std(axis=0)
    sd[sd < 1e-6] = 1.0
    return np.concatenate([mu, 1.0 / sd]).astype(np.float32)


def generate_samples(ctx, n_chains=8, seed=7, L_range=(12, 24), D_range=(16, 64), sweeps=2, iters=4, nsamp=1024,
                     sites_per_chain=6):
    """Samples from random XXZ chains (couplings as the batched workload):
    each chain is swept `sweeps` times (FP64 Lanczos) to a realistic state,
    then a handful of sites contribute exact Davidson corrections."""
    rng = np.random.default_rng(seed)
    Xs, As, Ts = [], [], []
    o = SweepOpts.default(precision=TN_FP64, env_precision=TN_FP64, refine=0, gauge=1, eig_tol=1e-9,
                          max_krylov=32, max_restarts=1)
    for k, (L, D, J, Jz, hz) in enumerate(batched_workload(n_chains, seed, L_range, D_range)):
        ch = Chain(ctx, xxz_mpo(L, 0.5, Jxy=J, Jz=Jz, hz=hz), D)
        ch.initialize(1, seed * 1000 + k)
        ch.normalize_qr(1)
        ch.load_caches(o)
        for _ in range(sweeps):
            ch.sweep(0, o)
            ch.sweep(1, o)
        # after a right-to-left pass the orthogonality centre is at site 0 and
        # every right environment is current; the sampled sites use those
        sites = sorted(set(int(s) for s in rng.integers(0, max(1, L // 3), sites_per_chain)))
        for s in sites:
            X, a, t = ch.precond_samples(s, iters=iters, nsamp=nsamp, seed=int(rng.integers(1 << 30)))
            Xs.append(X)
            As.append(a)
            Ts.append(t)
        ch.close()
    return np.concatenate(Xs), np.concatenate(As), np.concatenate(Ts)


def snapshot_rescale(g, a, t, nsamp):
    """Per snapshot s (nsamp consecutive samples): Davidson is blind to the
    scale of its correction, so the objective is the cosine between the
    correction g*a and the exact correction t,
        c_s = (sum g a t)^2 / (sum g^2 a^2 * sum t^2).
    Its gradient in g equals that of sum (g a' - t')^2 with t' = t/|t|_s and
    a' = k_s a, k_s = (sum g a t)/(sum g^2 a^2 |t|_s): returns (a', t', mean c)."""
    ns = len(a) // nsamp
    a2, t2, cs = np.empty_like(a), np.empty_like(t), []
    for s in range(ns):
        sl = slice(s * nsamp, (s + 1) * nsamp)
        ga, ts = g[sl] * a[sl], t[sl]
        A, B, E = np.dot(ga, ts), np.dot(ga, ga), np.dot(ts, ts)
        if B <= 0 or E <= 0:
            a2[sl], t2[sl] = 0.0, 0.0
            continue
        a2[sl] = a[sl] * (A / (B * np.sqrt(E)))
        t2[sl] = ts / np.sqrt(E)
        cs.append(A * A / (B * E))
    return a2, t2, float(np.mean(cs)) if cs else 0.0


def train(ctx, X, a, t, nsamp, params=None, norm=None, rounds=6, steps=400, batch=4096, lr=2e-3, seed=0):
    """Maximise the mean per-snapshot cosine between the learned correction and
    the exact one: alternate the per-snapshot rescale (network forward on the
    device) with Adam steps on the resulting squared loss (device kernels).
    Returns (params, norm, (cosine before, cosine after))."""
    if norm is None:
        norm = fit_norm(X)
    p = init_params(seed) if params is None else np.array(params, dtype=np.float32)
    c0 = None
    for r in range(rounds):
        g = ctx.mlp_forward(p, norm, X)
        a2, t2, c = snapshot_rescale(g, a, t, nsamp)
        if c0 is None:
            c0 = c
        p, _ = ctx.mlp_train(p, norm, X, a2, t2, steps, batch, lr, seed + r)
    g = ctx.mlp_forward(p, norm, X)
    _, _, c1 = snapshot_rescale(g, a, t, nsamp)
    return p, norm, (c0, c1)


def save(path, params, norm):
    np.savez(path, params=np.asarray(params, dtype=np.float32), norm=np.asarray(norm, dtype=np.float32))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), z["norm"].astype(np.float32)
