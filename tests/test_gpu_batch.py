"""Batched chains (tn_batch_*): many chains advanced in lockstep with grouped
launches must reproduce the one-chain sweep of each chain, and converge to
exact diagonalisation where the bond dimension is exact."""
import numpy as np
import pytest

import oracle as O
import tensornetworks_amd as tn
from tensornetworks_amd.models import xxz_mpo, batched_workload

pytestmark = pytest.mark.gpu


def fp32_opts(**kw):
    return tn.SweepOpts.default(precision=tn.TN_FP32, env_precision=tn.TN_FP32, refine=0, gauge=1,
                                eig_tol=kw.pop("eig_tol", 1e-9), max_krylov=kw.pop("max_krylov", 40),
                                max_restarts=kw.pop("max_restarts", 2), **kw)


def make_chain(ctx, L, D, J, Jz, hz, seed, opts):
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5, Jxy=J, Jz=Jz, hz=hz), D)
    ch.initialize(1, seed)
    ch.normalize_qr(1)
    ch.load_caches(opts)
    return ch


def test_batch_matches_single_chain_sweeps(ctx):
    # chains of different lengths and bond dimensions (rounds where some
    # chains are idle, edge sites, padded and unpadded shapes)
    spec = [(8, 8, 1.0, 1.0, 0.0), (12, 16, 0.7, 1.6, 0.3), (6, 32, 1.3, 0.4, -0.2), (10, 12, 0.9, 1.1, 0.45),
            (14, 40, 1.1, 0.8, -0.1)]
    o = fp32_opts()
    batch_chains = [make_chain(ctx, *s, seed=100 + k, opts=o) for k, s in enumerate(spec)]
    solo_chains = [make_chain(ctx, *s, seed=100 + k, opts=o) for k, s in enumerate(spec)]
    b = tn.Batch(ctx, batch_chains, max_krylov=40)
    st = tn.SweepStats()
    for _ in range(3):
        b.sweep(0, o, st)
        b.sweep(1, o, st)
    for ch in solo_chains:
        for _ in range(3):
            ch.sweep(0, o)
            ch.sweep(1, o)
    for (L, *_), cb, cs in zip(spec, batch_chains, solo_chains):
        eb, es = cb.expectation(), cs.expectation()
        assert abs(eb - es) < 1e-8 * max(1.0, abs(es)), (L, eb, es)
        assert cb.dims() == cs.dims()
    assert st.n_matvec > 0 and st.n_qr > 0
    # the bulk sites apply their identity environment channels as copies
    assert st.n_ident_left > 0 and st.n_ident_right > 0
    b.close()


def test_batch_partial_reorth_matches_single_chain_sweeps(ctx):
    # partial re-orthogonalisation in the pipelined batch (a chain whose new
    # vector needs a pass against the whole basis defers its tridiagonal
    # solve by one iteration) against the one-chain sweeps with the same
    # option, and against the batch with a pass every iteration
    spec = [(8, 8, 1.0, 1.0, 0.0), (12, 16, 0.7, 1.6, 0.3), (6, 32, 1.3, 0.4, -0.2), (10, 12, 0.9, 1.1, 0.45),
            (14, 40, 1.1, 0.8, -0.1)]
    energies, stats = {}, {}
    for reorth in (0, 1):
        o = fp32_opts(reorth=reorth)
        chains = [make_chain(ctx, *s, seed=100 + k, opts=o) for k, s in enumerate(spec)]
        b = tn.Batch(ctx, chains, max_krylov=40)
        st = tn.SweepStats()
        for _ in range(3):
            b.sweep(0, o, st)
            b.sweep(1, o, st)
        energies[reorth] = [ch.expectation() for ch in chains]
        stats[reorth] = st
        b.close()
    o = fp32_opts(reorth=1)
    solo = [make_chain(ctx, *s, seed=100 + k, opts=o) for k, s in enumerate(spec)]
    for ch in solo:
        for _ in range(3):
            ch.sweep(0, o)
            ch.sweep(1, o)
    for (L, *_), eb, ef, cs in zip(spec, energies[1], energies[0], solo):
        es = cs.expectation()
        assert abs(eb - es) < 1e-8 * max(1.0, abs(es)), (L, eb, es)
        assert abs(eb - ef) < 1e-8 * max(1.0, abs(ef)), (L, eb, ef)
    assert stats[0].n_full_reorth == stats[0].n_matvec
    assert stats[1].n_full_reorth < 0.5 * stats[1].n_matvec, (stats[1].n_full_reorth, stats[1].n_matvec)


def test_batch_identity_channels_same_sweeps(ctx):
    # the batched operator with identity channels copied gives the sweeps of
    # the full contraction
    spec = [(10, 16, 0.8, 1.2, 0.1), (7, 24, 1.2, 0.5, -0.3)]
    energies = []
    for skip in (0, 1):
        o = fp32_opts(skip_identity=skip)
        chains = [make_chain(ctx, *sp, seed=50 + k, opts=o) for k, sp in enumerate(spec)]
        b = tn.Batch(ctx, chains, max_krylov=40)
        st = tn.SweepStats()
        for _ in range(2):
            b.sweep(0, o, st)
            b.sweep(1, o, st)
        assert (st.n_ident_left > 0) == bool(skip)
        energies.append([ch.expectation() for ch in chains])
        b.close()
    for e0, e1 in zip(*energies):
        assert abs(e1 - e0) < 1e-6 * abs(e0), (e0, e1)


def test_batch_reaches_exact_diagonalisation(ctx):
    # D = 2^(L/2): the MPS manifold contains the exact ground state
    spec = [(8, 16, 1.2, 0.5, 0.2), (10, 32, 0.6, 1.8, -0.4), (6, 8, 1.0, 1.0, 0.0)]
    o = fp32_opts(eig_tol=1e-10)
    chains = [make_chain(ctx, *s, seed=7 + k, opts=o) for k, s in enumerate(spec)]
    b = tn.Batch(ctx, chains, max_krylov=40)
    for _ in range(6):
        b.sweep(0, o)
        b.sweep(1, o)
    for (L, D, J, Jz, hz), ch in zip(spec, chains):
        e_ed = O.ed_ground(L, 2, J, Jz, hz, 0.0, 1e-12, 400)
        e = ch.expectation()
        # FP32 contractions: the variational energy is within ~1e-6 relative
        assert abs(e - e_ed) < 2e-6 * abs(e_ed), (L, e, e_ed)
        assert e > e_ed - 1e-9 * abs(e_ed)  # variational bound (energy evaluated in FP64)
    b.close()


def test_batch_rejects_fp64_options(ctx):
    o = fp32_opts()
    ch = make_chain(ctx, 6, 8, 1.0, 1.0, 0.0, 1, o)
    b = tn.Batch(ctx, [ch], max_krylov=16)
    with pytest.raises(tn.TnError):
        b.sweep(0, tn.SweepOpts.default(precision=tn.TN_FP64, env_precision=tn.TN_FP64, max_krylov=16))
    b.close()


def test_batch_of_128_chains_matches_single_chain_sweeps(ctx):
    # the batched workload's chain count: 128 chains of the batched workload's
    # couplings (lengths 8-24, bond dimensions 8-48 to keep the test short),
    # one batched sweep pair against the same sweeps chain by chain from the
    # same states; converged local solves in FP32 contractions (whose
    # summation order differs between the grouped and the single launches)
    # agree to |dE|/L <= 1e-8
    o = fp32_opts(reorth=1)
    spec = batched_workload(128, 4242, L_range=(8, 24), D_range=(8, 48))
    batch_chains = [make_chain(ctx, *s, seed=500 + k, opts=o) for k, s in enumerate(spec)]
    b = tn.Batch(ctx, batch_chains, max_krylov=40)
    st = tn.SweepStats()
    b.sweep(0, o, st)
    b.sweep(1, o, st)
    eb = [c.expectation() for c in batch_chains]
    b.close()
    for c in batch_chains:
        c.close()
    worst = 0.0
    for k, s in enumerate(spec):
        c = make_chain(ctx, *s, seed=500 + k, opts=o)
        c.sweep(0, o)
        c.sweep(1, o)
        es = c.expectation()
        worst = max(worst, abs(eb[k] - es) / s[0])
        c.close()
    assert worst < 1e-8, worst
    # nearly every local solve converges (the counts cover all 128 chains)
    assert st.n_solve_converged > 0 and st.n_solve_capped <= 0.01 * st.n_solve_converged
