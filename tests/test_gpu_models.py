"""Whole sweeps on the GPU for the other models of the reference Factory
(transverse-field Ising, long-range, three-body), the S=1 chains of the
reference tests and the S=1 + Dz workload, against the reference's recorded
energies and exact diagonalisation; the compressed <H^2> MPO on the device."""
import json
import os

import numpy as np
import pytest

import oracle as O
import tensornetworks_amd as tn
from tensornetworks_amd.models import (xxz_mpo, tfim_mpo, long_range_mpo, three_body_mpo, mpo_square, identity_mpo)

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))


def model(case):
    m, L, S, f = case["model"], case["L"], case["S"], case["form"]
    if m == "tfim":
        return tfim_mpo(L, S, case["J"], case["hx"], form=f)
    if m == "three_body":
        return three_body_mpo(L, S, case["J"], case["K"], case["hz"], form=f)
    if m == "long_range":
        return long_range_mpo(L, S, case["J"], case["hx"], case["NN"], form=f)
    raise ValueError(m)


def maxdw(Ws):
    return max(max(W.shape[0], W.shape[1]) for W in Ws)


@pytest.mark.parametrize("case", GOLD["dmrg_models"], ids=lambda c: f"{c['model']}-L{c['L']}-f{c['form']}-"
                         f"D{c['schedule'][-1][1]}")
def test_model_ground_state_matches_reference(ctx, case):
    # VariationalGroundStateTests.C:426-638: Factory model -> (CanonicalForm +
    # Compress(Parker, 0, 1e-10) for the long-range and three-body chains) ->
    # FindVariationalGroundState -> E/(L-1) and <H^2> - <H>^2 with CreateH2Operator
    Ws = model(case)
    f = case["form"]
    if case.get("compress_first"):
        Ws, _ = tn.mpo_reduce(Ws, f, ops=3, eps=1e-10)
    L = case["L"]
    D0 = case["schedule"][0][1]
    ch = tn.Chain(ctx, Ws, D0, Dcap=max(s[1] for s in case["schedule"]))
    ch.initialize(1, 2024 + L)
    H2, _ = tn.mpo_reduce(mpo_square(Ws), f, ops=3, eps=1e-13)
    sched = [tn.ScheduleLine(mi, Dm, 1, eps_energy=1e-9, eps_lambda=1e-9, eps_eigen=1e-12) for mi, Dm in case["schedule"]]
    res = tn.FindVariationalGroundState(ch, sched, tn.SweepOpts.default(eig_tol=1e-12), H2=H2)
    E1 = ch.expectation()
    E2 = ch.expectation(H2)
    assert abs(E1 / (L - 1) - case["E"]) < case["tol"], (E1 / (L - 1), case["E"])
    assert abs(E2 - E1 * E1) < max(case["E2_minus_E1sq"], 1e-12), (E2, E1 * E1)
    assert res.nsweeps < case["schedule"][-1][0]
    # variational: not below exact diagonalisation
    assert E1 > O.ed_ground_mpo(Ws, tol=1e-13) - 1e-10
    ch.close()


def test_compressed_h2_expectation_equals_uncompressed(ctx):
    # the compression is exact: <H^2> through the Dw=25 product MPO and through
    # the compressed one agree on an arbitrary (unconverged) state
    for S, form, Dz in ((0.5, 0, 0.0), (1.0, 1, 0.1)):
        Ws = xxz_mpo(10, S, Dz=Dz, form=form)
        H2 = mpo_square(Ws)
        H2c, _ = tn.mpo_reduce(H2, form, ops=3, eps=1e-13)
        assert maxdw(H2c) < maxdw(H2)
        ch = tn.Chain(ctx, Ws, 16)
        ch.initialize(1, 31)
        ch.normalize(1)
        a, b = ch.expectation(H2), ch.expectation(H2c)
        assert abs(a - b) < 1e-12 * abs(a), (a, b)
        ch.close()


@pytest.mark.parametrize("case", GOLD["dmrg_per_bond_s1"], ids=lambda c: f"S1-L{c['L']}")
def test_s1_grow_d27_matches_reference(ctx, case):
    # VariationalGroundStateTests.C:299-372 (RunLongTests): S=1 Heisenberg, D
    # grown 2 -> 27 over the reference's schedule lines
    L = case["L"]
    Ws = xxz_mpo(L, 1.0, form=case["form"])
    ch = tn.Chain(ctx, Ws, 2, Dcap=27)
    ch.initialize(1, 4242 + L)
    sched = [tn.ScheduleLine(mi, Dm, dD, eps_energy=eE, eps_lambda=1e-12, eps_eigen=1e-12)
             for mi, Dm, dD, eE in case["schedule"]]
    tn.FindVariationalGroundState(ch, sched, tn.SweepOpts.default(eig_tol=1e-12))
    E = ch.expectation() / (L - 1)
    assert abs(E - case["E"]) < case["tol"], (E, case["E"])
    ch.close()


def _s1_ed_chain(ctx, prec, refine):
    # S=1 + Dz=0.1, L=8: D=81 = 3^4 holds every state, so the sweep reaches the
    # exact ground state of the 6561-dimensional space
    L, D = 8, 81
    Ws = xxz_mpo(L, 1.0, Dz=0.1)
    ch = tn.Chain(ctx, Ws, D)
    ch.initialize(1, 77)
    ch.normalize_qr(1)
    o = tn.SweepOpts.default(eig_tol=1e-11, gauge=1, precision=prec,
                             env_precision=(tn.TN_FP32 if prec in (tn.TN_BF16, tn.TN_BF16X3) else prec),
                             refine=refine, refine_precision=tn.TN_FP32 if prec == tn.TN_BF16 else tn.TN_FP64,
                             refine_tol=1e-10, max_krylov=40, reorth=1)
    ch.load_caches(o)
    for _ in range(6):
        ch.sweep(0, o)
        ch.sweep(1, o)
    return ch


@pytest.mark.parametrize("prec,refine,tol", [(tn.TN_FP64, 0, 1e-10), (tn.TN_FP32, 0, 1e-8), (tn.TN_FP32, 1, 1e-9),
                                             (tn.TN_BF16, 1, 1e-8), (tn.TN_BF16X3, 0, 1e-8),
                                             (tn.TN_BF16X3, 1, 1e-9)])
def test_s1_dz_sweep_matches_exact_diagonalisation(ctx, prec, refine, tol):
    L = 8
    ed = O.ed_ground(L, 3, Dz=0.1, tol=1e-13, maxit=400)
    assert abs(O.ed_ground_mpo(xxz_mpo(L, 1.0, Dz=0.1), tol=1e-13) - ed) < 1e-10
    ch = _s1_ed_chain(ctx, prec, refine)
    E = ch.expectation() / ch.expectation(identity_mpo(L, 3))
    assert abs(E - ed) / L < tol, (E, ed)
    assert E > ed - 1e-10
    ch.close()


def test_mixed_precision_matches_fp64_at_bulk_shapes(ctx, tmp_path):
    # the benched scheme (QR gauge, identity channels copied, partial
    # re-orthogonalisation, FP32 contractions; also BF16 contractions with the
    # FP32 refinement) on an S=1 + Dz chain at D=256 (3072 x 4096 x 256 stage-1
    # GEMMs in the bulk): from the same state, the mixed sweeps end within
    # |dE/L| <= 1e-8 of the FP64 sweeps
    L, D = 20, 256
    Ws = xxz_mpo(L, 1.0, Dz=0.1)
    base = dict(eig_tol=1e-8, max_krylov=24, max_restarts=2, gauge=1, skip_identity=1, reorth=1, refine=0)
    o64 = tn.SweepOpts.default(precision=tn.TN_FP64, env_precision=tn.TN_FP64, **base)
    ch = tn.Chain(ctx, Ws, D)
    ch.initialize(1, 20261202)
    ch.normalize_qr(1)
    ch.load_caches(o64)
    for _ in range(3):  # reach the converged regime first
        ch.sweep(0, o64)
        ch.sweep(1, o64)
    path = str(tmp_path / "warm.tnck")
    ch.save(path)
    D1, D2 = ch.dims()
    assert max(D2) == D
    ch.close()

    def run(opts):
        c, _ = tn.Chain.load(ctx, path)
        c.load_caches(opts)
        for _ in range(2):
            c.sweep(0, opts)
            c.sweep(1, opts)
        e = c.expectation() / c.expectation(identity_mpo(L, 3))
        c.close()
        return e

    e64 = run(o64)
    e32 = run(tn.SweepOpts.default(precision=tn.TN_FP32, env_precision=tn.TN_FP32, **base))
    assert abs(e32 - e64) / L <= 1e-8, (e32, e64)
    ob = dict(base, refine=1, refine_precision=tn.TN_FP32)
    ebf = run(tn.SweepOpts.default(precision=tn.TN_BF16, env_precision=tn.TN_FP32, **ob))
    assert abs(ebf - e64) / L <= 1e-8, (ebf, e64)
    # three-piece bfloat16 contractions (FP32 environment updates)
    ex3 = run(tn.SweepOpts.default(precision=tn.TN_BF16X3, env_precision=tn.TN_FP32, **base))
    assert abs(ex3 - e64) / L <= 1e-8, (ex3, e64)
