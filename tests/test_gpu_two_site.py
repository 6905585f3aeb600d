"""Two-site DMRG (the D-ramp workload, an extension of the reference's one-site
sweep) on the GPU vs the CPU oracle's two-site sweep from the same state, and
vs the one-site fixed point at equal D."""
import numpy as np
import pytest

import oracle as O
import tensornetworks_amd as tn
from tensornetworks_amd.models import xxz_mpo

pytestmark = pytest.mark.gpu


def test_two_site_matches_oracle_small(ctx):
    L, D0, seed = 12, 4, 99
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5), D0, Dcap=32)
    ch.initialize(1, seed)
    ch.normalize(1)
    o = tn.SweepOpts.default(eig_tol=1e-12)
    ch.load_caches(o)
    m = O.OracleMPS(L, 2, D0, init=1, seed=seed)
    m.normalize(1)
    m.load_caches()
    for Dmax in (8, 16, 16, 16):
        ch.sweep2(0, Dmax, 0.0, o)
        ch.sweep2(1, Dmax, 0.0, o)
        m.sweep2(0, Dmax, 0.0, 1e-12)
        m.sweep2(1, Dmax, 0.0, 1e-12)
    e_gpu, e_cpu = ch.expectation(), m.expectation(0)
    # SU(2) multiplets in the entanglement spectrum make the kept subspace at a
    # truncation boundary ambiguous, so agreement is at the truncation level
    assert abs(e_gpu - e_cpu) < 1e-6
    assert ch.dims() == m.dims()
    assert abs(e_gpu - O.ed_ground(L, 2)) < 1e-6  # D=16 at L=12 is close to exact


def test_two_site_ramp_reaches_one_site_value(ctx):
    # L=10 D=8: -0.473113399629092 per bond (PyTensorNetworks/test_Unittest1.py:22)
    L = 10
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5), 2, Dcap=8)
    ch.initialize(1, 5)
    ch.normalize(1)
    o = tn.SweepOpts.default(eig_tol=1e-12)
    ch.load_caches(o)
    for Dmax in (4, 8, 8, 8, 8, 8, 8, 8):
        st, tr = ch.sweep2(0, Dmax, 0.0, o)
        st, tr = ch.sweep2(1, Dmax, 0.0, o)
    assert abs(ch.expectation() / 9 - (-0.473113399629092)) < 1e-7


def test_two_site_eps_truncation(ctx):
    # SVCompressor with eps > 0: kept dimension follows the discarded weight
    L = 10
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5), 4, Dcap=32)
    ch.initialize(1, 6)
    ch.normalize(1)
    o = tn.SweepOpts.default(eig_tol=1e-12)
    ch.load_caches(o)
    for _ in range(4):
        ch.sweep2(0, 32, 1e-8, o)
        st, tr = ch.sweep2(1, 32, 1e-8, o)
    assert tr < 1e-8
    assert abs(ch.expectation() - O.ed_ground(L, 2)) < 1e-6


def test_two_site_L64_ramp_to_256(ctx):
    # the two-site workload at its full size: S=1/2 Heisenberg L=64, two-site
    # sweeps at Dmax 32, 64, 128, 256 (FP64, tolerance 1e-12, no cutoff).
    # What the mathematics fixes: the energy falls with Dmax, the truncation at
    # 256 is negligible, and one-site sweeps at D=256 from the final state
    # (which cannot leave the kept space) agree with it to the truncation level
    L = 64
    o = tn.SweepOpts.default(precision=tn.TN_FP64, env_precision=tn.TN_FP64, eig_tol=1e-12, max_krylov=24,
                             max_restarts=8, gauge=0, reorth=1)
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5), 32, Dcap=256)
    ch.initialize(1, 20261202)
    ch.normalize(1)
    ch.load_caches(o)
    es = []
    for Dmax in (32, 64, 128, 256):
        ch.sweep2(0, Dmax, 0.0, o)
        _, tr = ch.sweep2(1, Dmax, 0.0, o)
        es.append(ch.expectation() / (L - 1))
    assert max(ch.dims()[1]) == 256
    assert all(b < a + 1e-12 for a, b in zip(es, es[1:])), es
    assert tr < 1e-12, tr
    e2 = es[-1]
    ch.load_caches(o)
    ch.sweep(0, o)
    ch.sweep(1, o)
    e1 = ch.expectation() / (L - 1)
    assert abs(e1 - e2) < 1e-9, (e1, e2)
    ch.close()
