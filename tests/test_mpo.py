"""MPO models of the reference Factory and the MPO canonical form / compression
(tn_mpo_reduce, host code of the C ABI library; no device needed).

Pins:
* the compressed bond dimensions the reference's own tests assert
  (tests/golden/reference_values.json "mpo_reduce", each with its file:line);
* exactness: the reduced MPO contracts to the same operator (dense, small L);
* what the mathematics fixes: on every bond the reduced MPO keeps the two
  identity channels plus exactly the rank of the operator's part that acts
  non-trivially on both sides of the bond (operator Schmidt rank after
  projecting out the identity on either side);
* the model W tables: closed forms (free-fermion transverse-field Ising chain,
  classical long-range Ising by brute force) and explicit sums of local terms.
"""
import itertools
import json
import os

import numpy as np
import pytest

import oracle as O
from tensornetworks_amd import mpo_reduce, TnError
from tensornetworks_amd.models import (xxz_mpo, tfim_mpo, long_range_mpo, three_body_mpo, mpo_square, mpo_contract,
                                       spin_ops, sx_op)

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_values.json")))


def maxdw(Ws):
    return max(max(W.shape[0], W.shape[1]) for W in Ws)


def model(case):
    m, L, S, f = case["model"], case["L"], case["S"], case["form"]
    if m == "heisenberg":
        return xxz_mpo(L, S, form=f)
    if m == "tfim":
        return tfim_mpo(L, S, case["J"], case["hx"], form=f)
    if m == "three_body":
        return three_body_mpo(L, S, case["J"], case["K"], case["hz"], form=f)
    if m == "long_range":
        return long_range_mpo(L, S, case["J"], case["hx"], case["NN"], form=f)
    raise ValueError(m)


@pytest.mark.parametrize("case", GOLD["mpo_reduce"], ids=lambda c: f"{c['model']}-L{c['L']}-{c['op']}-f{c['form']}")
def test_reduced_dw_matches_reference(case):
    H = model(case)
    f = case["form"]
    if case["op"] == "H":
        Wc, _ = mpo_reduce(H, f, ops=1)
        assert maxdw(Wc) == case["dw_canonical"]
        Wr, te = mpo_reduce(H, f, ops=3, eps=case["eps"])
        target = H
    else:
        if case["op"] == "H2 of compressed H":
            H, _ = mpo_reduce(H, f, ops=3, eps=1e-10)
        target = mpo_square(H)
        # Hamiltonian::CreateH2Operator: product, CanonicalForm, Compress(Parker, 0, 1e-13)
        Wr, te = mpo_reduce(target, f, ops=3, eps=1e-13)
    assert maxdw(Wr) == case["dw"]
    assert te < 1e-8
    A, B = mpo_contract(target), mpo_contract(Wr)
    assert np.abs(A - B).max() < 1e-12 * max(1.0, np.abs(A).max())


def interior_rank(Op, d, L, k, tol=1e-10):
    """Rank of the part of Op (d^L x d^L, site 0 most significant) acting
    non-trivially on both sides of the bond after site k-1 (k sites left)."""
    nl, nr = d ** k, d ** (L - k)
    T = Op.reshape(nl, nr, nl, nr).transpose(0, 2, 1, 3).reshape(nl * nl, nr * nr)
    il = np.eye(nl).ravel() / np.sqrt(nl)
    ir = np.eye(nr).ravel() / np.sqrt(nr)
    T = T - np.outer(il, il @ T)
    T = T - np.outer(T @ ir, ir)
    s = np.linalg.svd(T, compute_uv=False)
    return int((s > tol * max(s[0], 1.0)).sum()) if s.size else 0


@pytest.mark.parametrize("name,Ws,form", [
    ("heisenberg L6 H2", mpo_square(xxz_mpo(6, 0.5)), 0),
    ("S=1 + Dz L5 H2", mpo_square(xxz_mpo(5, 1.0, Dz=0.1, hz=0.2)), 0),
    ("tfim L6 H2 lower", mpo_square(tfim_mpo(6, 0.5, 1.0, 0.7, form=1)), 1),
    ("long-range L7 NN3", long_range_mpo(7, 0.5, 1.0, 0.4, 3), 0),
    ("three-body long-range L7 NN2", long_range_mpo(7, 0.5, 0.8, 0.3, 2, True, form=1), 1),
])
def test_bond_dims_are_identity_channels_plus_interior_rank(name, Ws, form):
    Wr, _ = mpo_reduce(Ws, form, ops=3, eps=1e-13)
    Op = mpo_contract(Ws)
    d, L = Ws[0].shape[2], len(Ws)
    for k in range(1, L):
        assert Wr[k - 1].shape[1] == 2 + interior_rank(Op, d, L, k), (name, k)
    assert np.abs(mpo_contract(Wr) - Op).max() < 1e-11 * max(1.0, np.abs(Op).max())


def test_canonical_form_alone_is_exact_and_upper_equals_lower():
    for f in (0, 1):
        Ws = mpo_square(xxz_mpo(5, 1.0, Dz=0.1, form=f))
        Wc, te = mpo_reduce(Ws, f, ops=1)
        assert te == 0.0
        assert np.abs(mpo_contract(Wc) - mpo_contract(Ws)).max() < 1e-12
    a, _ = mpo_reduce(mpo_square(xxz_mpo(7, 0.5, form=0)), 0)
    b, _ = mpo_reduce(mpo_square(xxz_mpo(7, 0.5, form=1)), 1)
    assert [W.shape for W in a] == [W.shape for W in b]


def test_reduce_rejects_irregular_mpo():
    Ws = xxz_mpo(4, 0.5)
    bad = [W.copy() for W in Ws]
    bad[1][1, 0] = np.eye(2)  # identity leaks into the first column below row 0
    with pytest.raises(TnError):
        mpo_reduce(bad, 0)


def tfim_exact(L, J, hx):
    """Open chain H = J sum Sz Sz + hx sum Sx (S=1/2): free fermions,
    E0 = -sum of singular values of the bidiagonal matrix diag(hx/2), superdiag J/4."""
    B = np.diag([hx / 2.0] * L) + np.diag([J / 4.0] * (L - 1), 1)
    return -np.linalg.svd(B, compute_uv=False).sum()


@pytest.mark.parametrize("L,J,hx,form", [(6, 1.0, 1.0, 0), (9, 1.0, 1.0, 1), (8, -0.7, 0.3, 0), (10, 1.0, 2.0, 1)])
def test_tfim_mpo_free_fermion_energy(L, J, hx, form):
    Ws = tfim_mpo(L, 0.5, J, hx, form)
    assert abs(O.ed_ground_mpo(Ws, tol=1e-13) - tfim_exact(L, J, hx)) < 1e-10
    if L <= 8:
        assert abs(np.linalg.eigvalsh(mpo_contract(Ws))[0] - tfim_exact(L, J, hx)) < 1e-10


def test_long_range_ising_brute_force():
    # hx = 0: diagonal; ground energy = min over classical configurations
    L, NN, J = 9, 3, 1.0
    best = min(sum(J / (r * r) * c[k] * c[k + r] for r in range(1, NN + 1) for k in range(L - r))
               for c in itertools.product([0.5, -0.5], repeat=L))
    for f in (0, 1):
        Ws = long_range_mpo(L, 0.5, J, 0.0, NN, form=f)
        assert abs(O.ed_ground_mpo(Ws, tol=1e-13) - best) < 1e-10
    # VariationalGroundStateTests.C:602 (D=2 reaches this classical state)
    case = [c for c in GOLD["dmrg_models"] if c["model"] == "long_range"][0]
    assert abs(best / (L - 1) - case["E"]) < case["tol"]


def kron_chain(ops, L, d):
    """Operator with ops {site: matrix} on a chain of L sites, identity elsewhere."""
    out = np.eye(1)
    for i in range(L):
        out = np.kron(out, ops.get(i, np.eye(d)))
    return out


@pytest.mark.parametrize("three_body", [False, True])
def test_long_range_mpo_is_sum_of_local_terms(three_body):
    L, NN, J, hx, d = 6, 3, 0.8, 0.35, 2
    _, _, Sz = spin_ops(d)
    Sx = sx_op(d)
    H = sum(kron_chain({k: hx * Sx}, L, d) for k in range(L))
    for r in range(1, NN + 1):
        for k in range(L - r):
            if three_body:
                ops = {k: Sz @ Sz, k + r: J / (r * r) * Sz @ Sz}
                for j in range(1, r):
                    ops[k + j] = Sz / (j * j)
            else:
                ops = {k: Sz, k + r: J / (r * r) * Sz}
            H = H + kron_chain(ops, L, d)
    for f in (0, 1):
        assert np.abs(mpo_contract(long_range_mpo(L, 0.5, J, hx, NN, three_body, form=f)) - H).max() < 1e-13


def test_three_body_mpo_acts_as_sx_sx_plus_field():
    # channels 2 and 3 of Hamiltonian_1D_3Body.C:26-56 are linked only to each
    # other: the operator is J sum Sx Sx + hz sum Sz, and compression finds Dw = 3
    L, d, J, K, hz = 6, 2, 0.9, 1.3, 0.4
    _, _, Sz = spin_ops(d)
    Sx = sx_op(d)
    H = sum(kron_chain({k: hz * Sz}, L, d) for k in range(L))
    H = H + sum(kron_chain({k: J * Sx, k + 1: Sx}, L, d) for k in range(L - 1))
    for f in (0, 1):
        Ws = three_body_mpo(L, 0.5, J, K, hz, form=f)
        assert np.abs(mpo_contract(Ws) - H).max() < 1e-13
    # the reference's golden at L=9 (D=4) lies above the exact energy, by less than 1e-5 per bond
    case = [c for c in GOLD["dmrg_models"] if c["model"] == "three_body"][0]
    E = O.ed_ground_mpo(three_body_mpo(9, 0.5, case["J"], case["K"], case["hz"]), tol=1e-13) / 8
    assert E < case["E"] < E + 1e-5


def test_ed_mpo_matches_xxz_ed():
    for L, d, Dz, hz in ((10, 2, 0.0, 0.0), (8, 3, 0.1, 0.0), (7, 3, -0.2, 0.3)):
        a = O.ed_ground_mpo(xxz_mpo(L, 0.5 * (d - 1), hz=hz, Dz=Dz), tol=1e-13)
        b = O.ed_ground(L, d, hz=hz, Dz=Dz, tol=1e-13)
        assert abs(a - b) < 1e-10
