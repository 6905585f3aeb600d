"""Failure handling, forced from the host: non-finite entries in a site
tensor, Lanczos breakdown on an exhausted Krylov space, local solves that
cannot converge, a preconditioner network with non-finite parameters.  Each
ends in an error code or in a documented fallback; the kernels treat the
non-finite values as data (no data-dependent indexing), so nothing faults."""
import numpy as np
import pytest

import oracle as O
import tensornetworks_amd as tn
from tensornetworks_amd import precond
from tensornetworks_amd.models import xxz_mpo

pytestmark = pytest.mark.gpu

# PyTensorNetworks/test_Unittest1.py:22 : L=10 S=1/2 D=8 -> -0.473113399629092 per bond
E_REF = -0.473113399629092


def reference_chain(ctx, opts):
    ch = tn.Chain(ctx, xxz_mpo(10, 0.5), 8)
    ch.initialize(1, 20261202)
    ch.normalize(1)
    ch.load_caches(opts)
    return ch


def test_non_finite_site_raises_and_context_recovers(ctx):
    o = tn.SweepOpts.default(eig_tol=1e-12)
    ch = reference_chain(ctx, o)
    D1, D2 = ch.dims()
    M = ch.site(4)
    M[3] = np.nan
    ch.set_site(4, D1[4], D2[4], M)
    with pytest.raises(tn.TnError):
        ch.load_caches(o)
        for _ in range(2):
            ch.sweep(0, o)
            ch.sweep(1, o)
    ch.close()
    # the context serves a fresh chain as before
    ch = reference_chain(ctx, o)
    for _ in range(10):
        ch.sweep(0, o)
        ch.sweep(1, o)
    assert abs(ch.expectation() / 9 - E_REF) < 1e-11


@pytest.mark.parametrize("reorth", [0, 1])
def test_lanczos_breakdown_on_exhausted_space(ctx, reorth):
    # L=4 at D=4 is exact and every local space has at most 16 elements, fewer
    # than the Krylov basis: each solve ends by breakdown (beta = 0 on an
    # invariant subspace) and the sweeps reach exact diagonalisation
    o = tn.SweepOpts.default(eig_tol=1e-13, max_krylov=64, reorth=reorth)
    ch = tn.Chain(ctx, xxz_mpo(4, 0.5), 4)
    ch.initialize(1, 3)
    ch.normalize(1)
    ch.load_caches(o)
    for _ in range(4):
        ch.sweep(0, o)
        ch.sweep(1, o)
    assert abs(ch.expectation() - O.ed_ground(4, 2)) < 1e-12


def test_unconverged_local_solves_stay_variational(ctx):
    # two Krylov vectors and no restart cannot reach 1e-14: the sweep goes on
    # with the best Ritz vector (the reference leaves an unconverged solve the
    # same way), and the energy stays finite and above the exact one
    o = tn.SweepOpts.default(eig_tol=1e-14, max_krylov=2, max_restarts=0)
    ch = reference_chain(ctx, o)
    for _ in range(3):
        ch.sweep(0, o)
        ch.sweep(1, o)
    e = ch.expectation()
    assert np.isfinite(e)
    assert e > -4.258035207282882  # ExactDiagTests.C:265-270
    assert e / 9 > E_REF + 1e-9  # not converged to the D=8 value


@pytest.mark.parametrize("kind", [2, 3])
def test_non_finite_network_gives_the_diagonal_correction(ctx, restore_mlp, kind):
    # a network whose output is NaN: the factor clamp turns it into a finite
    # constant, i.e. the diagonal correction up to scale, and the solves
    # converge to the reference energy
    ctx.set_mlp(np.full(tn.MLP_NP, np.nan, dtype=np.float32), np.array([0, 0, 0, 0, 1, 1, 1, 1], dtype=np.float32))
    o = tn.SweepOpts.default(eig_tol=1e-12, eig_method=1, precond=kind, max_krylov=40, max_restarts=4)
    ch = reference_chain(ctx, o)
    st = tn.SweepStats()
    for _ in range(10):
        ch.sweep(0, o, st)
        ch.sweep(1, o, st)
    assert abs(ch.expectation() / 9 - E_REF) < 1e-11
