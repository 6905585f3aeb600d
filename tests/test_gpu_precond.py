"""Learned Davidson preconditioner: the network's GPU kernels against the CPU
oracle (forward values, loss gradient), device training, and the Davidson
local solver against the reference energies."""
import numpy as np
import pytest

import oracle as O
import tensornetworks_amd as tn
from tensornetworks_amd import precond
from tensornetworks_amd.models import xxz_mpo

pytestmark = pytest.mark.gpu


def net(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(tn.MLP_NP) * scale).astype(np.float32)
    norm = np.concatenate([rng.standard_normal(4) * 0.2, 0.5 + rng.random(4)]).astype(np.float32)
    return p, norm


def test_mlp_forward_matches_oracle(ctx):
    p, norm = net(0)
    X = np.random.default_rng(1).standard_normal((5000, 4)) * 2.0
    g = ctx.mlp_forward(p, norm, X)
    ref = O.mlp_forward(p.astype(np.float64), norm.astype(np.float64), X)
    # FP32 network vs FP64 oracle
    assert np.abs(g - ref).max() < 2e-5 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [1, 255, 256, 1000, 4096])
def test_mlp_gradient_matches_oracle(ctx, n):
    p, norm = net(2)
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, 4))
    a = rng.standard_normal(n)
    t = rng.standard_normal(n)
    loss, g = ctx.mlp_loss_grad(p, norm, X, a, t)
    lref, gref = O.mlp_loss_grad(p.astype(np.float64), norm.astype(np.float64), X, a, t)
    assert loss == pytest.approx(lref, rel=1e-4)
    # FP32 kernels vs FP64 oracle (itself pinned by finite differences)
    assert np.linalg.norm(g - gref) < 1e-4 * np.linalg.norm(gref)


def test_mlp_training_reduces_loss(ctx):
    # targets from a smooth factor of the features: the device Adam fits it
    rng = np.random.default_rng(3)
    n = 20000
    X = rng.standard_normal((n, 4))
    a = rng.standard_normal(n)
    gtrue = 1.0 + 0.5 * np.tanh(X[:, 0]) - 0.3 * X[:, 2] / (1.0 + X[:, 2] ** 2)
    t = gtrue * a
    p0 = precond.init_params(0)
    norm = precond.fit_norm(X)
    p, (l0, l1) = ctx.mlp_train(p0, norm, X, a, t, steps=600, batch=2048, lr=3e-3, seed=1)
    assert l1 < 0.05 * l0, (l0, l1)
    # the reported loss is the oracle's loss of the returned parameters
    lref, _ = O.mlp_loss_grad(p.astype(np.float64), norm.astype(np.float64), X, a, t)
    assert l1 == pytest.approx(lref, rel=1e-3)


@pytest.mark.parametrize("kind", [0, 1])
def test_davidson_reference_energy(ctx, kind):
    # PyTensorNetworks/test_Unittest1.py:22 : L=10 S=1/2 D=8 -> -0.473113399629092 per bond
    ch = tn.Chain(ctx, xxz_mpo(10, 0.5), 8)
    ch.initialize(1, 20261202)
    ch.normalize(1)
    o = tn.SweepOpts.default(eig_tol=1e-12, eig_method=1, precond=kind, max_krylov=40, max_restarts=4)
    ch.load_caches(o)
    for _ in range(10):
        ch.sweep(0, o)
        ch.sweep(1, o)
    assert abs(ch.expectation() / 9 - (-0.473113399629092)) < 1e-11


def test_identity_network_equals_diagonal_davidson(ctx, restore_mlp):
    # w3 = 0, b3 = 1 makes the learned correction the diagonal one exactly;
    # one basis large enough that no solve restarts or falls back
    p = precond.init_params(5)
    norm = np.array([0, 0, 0, 0, 1, 1, 1, 1], dtype=np.float32)
    ctx.set_mlp(p, norm)
    res = {}
    for kind in (1, 2):
        ch = tn.Chain(ctx, xxz_mpo(12, 0.5, Jz=0.7, hz=0.1), 16)
        ch.initialize(1, 99)
        ch.normalize_qr(1)
        o = tn.SweepOpts.default(eig_tol=1e-10, eig_method=1, precond=kind, max_krylov=60, max_restarts=0, gauge=1)
        ch.load_caches(o)
        st = tn.SweepStats()
        for _ in range(2):
            ch.sweep(0, o, st)
            ch.sweep(1, o, st)
        res[kind] = (st.n_matvec, ch.expectation())
        assert st.n_precond_fallback == 0
    assert res[1][0] == res[2][0]
    assert res[1][1] == pytest.approx(res[2][1], abs=1e-12)


def test_precond_samples_are_exact_corrections(ctx):
    # Davidson's first iterate is u = x/|x| (x the site tensor), theta = u.Hu,
    # r = Hu - theta u.  Every sample of that iterate must be one component of
    # a = -r/(diag - theta)/|r| (|diag - theta| clamped to delta) and of the
    # exact correction t = (u0/(u.u0) - u)/|r|, u0 the local ground state, so
    # that u + |r| t is u0 up to normalisation.  Checked against the same
    # site solved on the host: environments zipped by the oracle from the
    # chain's own tensors, dense Heff (MPSSite_Contractions.C:155-193), numpy
    # eigh for u0
    L, d, s, delta = 8, 2, 3, 1e-2
    Ws = xxz_mpo(L, 0.5, Jz=1.3)
    ch = tn.Chain(ctx, Ws, 6)
    ch.initialize(1, 5)
    ch.normalize_qr(1)
    o = tn.SweepOpts.default(eig_tol=1e-10, gauge=1)
    ch.load_caches(o)
    X, a, t = ch.precond_samples(s, iters=1, nsamp=400, seed=1, delta=delta)
    assert X.shape == (400, 4) and np.all(np.isfinite(X)) and np.all(np.isfinite(a)) and np.all(np.isfinite(t))
    D1, D2 = ch.dims()
    F = np.ones(1)
    for i in range(s):
        F = O.env_left_ref(Ws[i], F, ch.site(i), d, D1[i], D2[i])
    G = np.ones(1)
    for i in reversed(range(s + 1, L)):
        G = O.env_right_ref(Ws[i], G, ch.site(i), d, D1[i], D2[i])
    H = O.heff_dense(Ws[s], F, G, d, D1[s], D2[s])
    assert H.shape == (72, 72) and np.abs(H - H.T).max() < 1e-12 * np.abs(H).max()
    w, Z = np.linalg.eigh(H)
    u0 = Z[:, 0]
    u = ch.site(s) / np.linalg.norm(ch.site(s))
    theta = u @ H @ u
    r = H @ u - theta * u
    nr = np.linalg.norm(r)
    den = np.diag(H) - theta
    pa = -r / np.where(np.abs(den) < delta, delta, den) / nr
    pt = (u0 / (u @ u0) - u) / nr
    px = np.arcsinh(den / delta)
    sa, st = np.abs(pa).max(), np.abs(pt).max()
    for q in range(len(a)):
        e = np.argmin(np.abs(pa - a[q]) / sa + np.abs(pt - t[q]) / st)
        assert abs(pa[e] - a[q]) < 1e-9 * sa and abs(pt[e] - t[q]) < 1e-8 * st, (q, e)
        assert abs(px[e] - X[q, 0]) < 1e-8 * max(1.0, abs(px[e]))
    # the samples reach many distinct components
    assert len(np.unique(np.round(t / st, 9))) > 40


@pytest.mark.parametrize("kind", [2, 3])
def test_learned_davidson_falls_back_to_lanczos(ctx, restore_mlp, kind):
    # one Davidson basis of 6 vectors cannot reach 1e-12: every unconverged
    # learned-correction solve is finished by plain Lanczos, counted in the
    # stats, and the sweeps still reach the reference energy
    # (PyTensorNetworks/test_Unittest1.py:22, L=10 S=1/2 D=8)
    ctx.set_mlp(precond.init_params(5), np.array([0, 0, 0, 0, 1, 1, 1, 1], dtype=np.float32))
    ch = tn.Chain(ctx, xxz_mpo(10, 0.5), 8)
    ch.initialize(1, 20261202)
    ch.normalize(1)
    o = tn.SweepOpts.default(eig_tol=1e-12, eig_method=1, precond=kind, max_krylov=6, max_restarts=40)
    ch.load_caches(o)
    st = tn.SweepStats()
    for _ in range(10):
        ch.sweep(0, o, st)
        ch.sweep(1, o, st)
    assert st.n_precond_fallback > 0
    assert abs(ch.expectation() / 9 - (-0.473113399629092)) < 1e-11


def test_davidson_rejects_unknown_correction(ctx):
    ch = tn.Chain(ctx, xxz_mpo(6, 0.5), 4)
    ch.initialize(1, 1)
    ch.normalize(1)
    o = tn.SweepOpts.default(eig_method=1, precond=4)
    ch.load_caches(o)
    with pytest.raises(tn.TnError):
        ch.sweep(0, o)


def test_trained_network_gains_on_held_out_corrections(ctx):
    # what training buys, on chains it has not seen: the mean per-snapshot
    # cosine between the correction and the exact one (the objective; Davidson
    # is blind to the correction's scale) rises well above the diagonal
    # correction's, which the untrained network reproduces exactly
    X, a, t = precond.generate_samples(ctx, n_chains=12, seed=11, L_range=(12, 24), D_range=(16, 64), iters=1,
                                       nsamp=1024, sites_per_chain=6)
    params, norm, (c0, c1) = precond.train(ctx, X, a, t, 1024, rounds=6, steps=300, batch=4096, lr=2e-3)
    assert c1 > c0
    Xh, ah, th = precond.generate_samples(ctx, n_chains=4, seed=991, L_range=(12, 24), D_range=(16, 64), iters=1,
                                          nsamp=1024, sites_per_chain=6)
    g0 = ctx.mlp_forward(precond.init_params(0), norm, Xh)
    assert np.abs(g0 - 1.0).max() < 1e-6
    _, _, cd = precond.snapshot_rescale(g0, ah, th, 1024)
    _, _, cl = precond.snapshot_rescale(ctx.mlp_forward(params, norm, Xh), ah, th, 1024)
    assert cl > cd + 0.25, (cd, cl)
