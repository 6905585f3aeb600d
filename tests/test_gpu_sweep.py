"""End-to-end one-site sweep on the GPU vs the reference's golden energies
and vs the CPU oracle run from the identical seeded state."""
import json
import os

import numpy as np
import pytest

import oracle as O
import tensornetworks_amd as tn
from tensornetworks_amd.models import xxz_mpo, identity_mpo

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))


def run_fixed_sweeps(ctx, L, S, D, nsweeps, seed, opts=None, Dz=0.0):
    Ws = xxz_mpo(L, S, Dz=Dz)
    ch = tn.Chain(ctx, Ws, D)
    ch.initialize(1, seed)
    ch.normalize(1)
    opts = opts or tn.SweepOpts.default(eig_tol=1e-12)
    ch.load_caches(opts)
    for _ in range(nsweeps):
        ch.sweep(0, opts)
        ch.sweep(1, opts)
    return ch


def test_L10_D8_ten_sweeps_matches_reference_value(ctx):
    # PyTensorNetworks/test_Unittest1.py:22 : L=10 S=1/2 D=8 -> -0.473113399629092 per bond
    ch = run_fixed_sweeps(ctx, 10, 0.5, 8, 10, 20261202)
    E = ch.expectation()
    assert abs(E / 9 - (-0.473113399629092)) < 1e-11
    # variational: above exact diagonalisation (ExactDiagTests.C:265-270)
    assert E > -4.258035207282882
    assert ch.expectation(identity_mpo(10, 2)) == pytest.approx(1.0, abs=1e-12)


def test_same_seed_same_energy_as_oracle(ctx):
    L, D, seed = 8, 6, 7
    ch = run_fixed_sweeps(ctx, L, 0.5, D, 3, seed)
    m = O.OracleMPS(L, 2, D, init=1, seed=seed)
    m.normalize(1)
    m.load_caches()
    for _ in range(3):
        m.sweep(0, 1e-12, 1)
        m.sweep(1, 1e-12, 1)
    assert abs(ch.expectation() - m.expectation(0)) < 1e-10
    assert ch.dims() == m.dims()


def test_normalisation_makes_right_canonical(ctx):
    L, D = 9, 5
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5), D)
    ch.initialize(1, 3)
    ch.normalize(1)
    D1, D2 = ch.dims()
    for i in range(1, L):
        M = ch.site(i).reshape(D2[i], D1[i], 2)  # [i2][i1][m]
        Bn = np.einsum("bam,bcm->ac", M, M)       # sum_m B B^T
        assert np.abs(Bn - np.eye(D1[i])).max() < 1e-12
    # bond dimensions min(D, d^k) from the right edge (MPSSite_Normalization.C:45-78)
    assert D1[L - 1] == 2 and D1[L - 2] == 4


@pytest.mark.parametrize("case", GOLD["dmrg_per_bond"], ids=lambda c: f"L{c['L']}-S{c['S']}-D{c['D']}")
def test_variational_ground_state_schedule(ctx, case):
    # every recorded one-site DMRG energy: S=1/2 up to L=19, S=1, S=5/2
    L, D, Dstart = case["L"], case["D"], case["Dstart"]
    Ws = xxz_mpo(L, case["S"])
    ch = tn.Chain(ctx, Ws, Dstart, Dcap=max(D, Dstart))
    ch.initialize(1, 1234 + L)
    sched = [tn.ScheduleLine(100, D, 1, eps_energy=1e-11, eps_lambda=1e-9, eps_eigen=1e-12)]
    res = tn.FindVariationalGroundState(ch, sched, tn.SweepOpts.default(eig_tol=1e-12))
    E = ch.expectation() / (L - 1)
    assert abs(E - case["E"]) < max(case["tol"], 1e-10), (case, E)
    if case["var"] is not None:
        assert res.de2 < case["var"]


def test_qr_gauge_same_fixed_point(ctx):
    # NormalizeQR gauge moves (MPSSite_Normalization.C:80-107) reach the same
    # variational minimum; bond values are cleared as in Bond::ClearSingularValues
    L, D = 10, 8
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5), D)
    ch.initialize(1, 20261202)
    ch.normalize_qr(1)
    opts = tn.SweepOpts.default(eig_tol=1e-12, gauge=1)
    ch.load_caches(opts)
    st = tn.SweepStats()
    for _ in range(10):
        ch.sweep(0, opts, st)
        ch.sweep(1, opts, st)
    assert abs(ch.expectation() / 9 - (-0.473113399629092)) < 1e-11
    assert st.n_qr > 0
    assert np.all(ch.bond_sv(4) == 0.0)


def test_qr_normalisation_right_canonical_large(ctx):
    L, D = 12, 96
    ch = tn.Chain(ctx, xxz_mpo(L, 1.0, Dz=0.1), D)
    ch.initialize(1, 11)
    ch.normalize_qr(1)
    D1, D2 = ch.dims()
    for i in range(1, L):
        M = ch.site(i).reshape(D2[i], D1[i], 3)
        Bn = np.einsum("bam,bcm->ac", M, M)
        assert np.abs(Bn - np.eye(D1[i])).max() < 1e-12, i
    # same state as the SVD normalisation up to the gauge: equal energies
    ch2 = tn.Chain(ctx, xxz_mpo(L, 1.0, Dz=0.1), D)
    ch2.initialize(1, 11)
    ch2.normalize(1)
    e1 = ch.expectation() / ch.expectation(identity_mpo(L, 3))
    e2 = ch2.expectation() / ch2.expectation(identity_mpo(L, 3))
    assert abs(e1 - e2) < 1e-10 * max(1.0, abs(e2))


@pytest.mark.parametrize("prec,tol", [(tn.TN_FP64, 1e-12), (tn.TN_FP32, 1e-6)])
def test_identity_channels_same_sweeps(ctx, prec, tol):
    # orthonormal blocks make the first/last MPO channel of each environment
    # the identity: applying it as a copy gives the sweeps of the full
    # contraction (FP64: to rounding; FP32: to the FP32 operator's accuracy)
    L, D = 16, 48
    energies, stats = [], []
    for skip in (0, 1):
        ch = tn.Chain(ctx, xxz_mpo(L, 1.0, Dz=0.1), D)
        ch.initialize(1, 99)
        ch.normalize_qr(1)
        o = tn.SweepOpts.default(eig_tol=1e-10, gauge=1, precision=prec, env_precision=prec, refine=0,
                                 skip_identity=skip)
        ch.load_caches(o)
        st = tn.SweepStats()
        for _ in range(3):
            ch.sweep(0, o, st)
            ch.sweep(1, o, st)
        energies.append(ch.expectation())
        stats.append(st)
        ch.close()
    assert abs(energies[1] - energies[0]) < tol * abs(energies[0]), energies
    assert stats[0].n_ident_left == 0 and stats[0].n_ident_right == 0
    # every bulk solve sees an identity channel on both sides
    solves = 6 * L
    assert stats[1].n_ident_left >= solves - 12 and stats[1].n_ident_right >= solves - 12, \
        (stats[1].n_ident_left, stats[1].n_ident_right)


def test_partial_reorth_matches_reference_value(ctx):
    # PyTensorNetworks/test_Unittest1.py:22 with partial re-orthogonalisation
    st = tn.SweepStats()
    o = tn.SweepOpts.default(eig_tol=1e-12, reorth=1)
    ch = tn.Chain(ctx, xxz_mpo(10, 0.5), 8)
    ch.initialize(1, 20261202)
    ch.normalize(1)
    ch.load_caches(o)
    for _ in range(10):
        ch.sweep(0, o, st)
        ch.sweep(1, o, st)
    assert abs(ch.expectation() / 9 - (-0.473113399629092)) < 1e-11
    assert st.n_full_reorth < st.n_matvec


@pytest.mark.parametrize("prec,tol", [(tn.TN_FP64, 1e-12), (tn.TN_FP32, 1e-6)])
def test_partial_reorth_same_sweeps(ctx, prec, tol):
    # semi-orthogonality at sqrt(unit roundoff) of the operator gives the
    # sweeps of full re-orthogonalisation (FP64: to rounding; FP32: to the
    # FP32 operator's accuracy) with far fewer passes against the whole basis
    L, D = 16, 48
    energies, stats = [], []
    for reorth in (0, 1):
        ch = tn.Chain(ctx, xxz_mpo(L, 1.0, Dz=0.1), D)
        ch.initialize(1, 99)
        ch.normalize_qr(1)
        o = tn.SweepOpts.default(eig_tol=1e-10, gauge=1, precision=prec, env_precision=prec, refine=0,
                                 max_krylov=40, reorth=reorth)
        ch.load_caches(o)
        st = tn.SweepStats()
        for _ in range(3):
            ch.sweep(0, o, st)
            ch.sweep(1, o, st)
        energies.append(ch.expectation())
        stats.append(st)
        ch.close()
    assert abs(energies[1] - energies[0]) < tol * abs(energies[0]), energies
    # every Lanczos iteration of the full mode has a pass against the whole basis
    assert stats[0].n_full_reorth == stats[0].n_matvec
    assert stats[1].n_full_reorth < 0.5 * stats[1].n_matvec, (stats[1].n_full_reorth, stats[1].n_matvec)
    assert abs(stats[1].n_matvec - stats[0].n_matvec) <= 0.1 * stats[0].n_matvec


def test_partial_reorth_two_site(ctx):
    # L=10 D=8 two-site ramp with partial re-orthogonalisation
    L = 10
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5), 2, Dcap=8)
    ch.initialize(1, 5)
    ch.normalize(1)
    o = tn.SweepOpts.default(eig_tol=1e-12, reorth=1)
    ch.load_caches(o)
    for Dmax in (4, 8, 8, 8, 8, 8, 8, 8):
        ch.sweep2(0, Dmax, 0.0, o)
        ch.sweep2(1, Dmax, 0.0, o)
    assert abs(ch.expectation() / 9 - (-0.473113399629092)) < 1e-7
