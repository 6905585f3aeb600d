"""Variational ground-state driver over the device chain.

Mirrors MPSImp::FindVariationalGroundState (TensorNetworksImp/MPS/MPSImp_Variational.C:16-62):
right-normalise, load the environment caches, then for every schedule line grow
D from the current maximum to Dmax in steps of DeltaD, sweeping (left pass +
right pass) until max|dE| < eps_E and max|d lambda| < eps_lambda or maxiter is
reached, and evaluate <H> and <H^2>.  Returns <H^2> - <H>^2 of the last line.
"""
from dataclasses import dataclass, field

from ._abi import SweepOpts, SweepStats, mpo_reduce
from .models import mpo_square, regular_form


@dataclass
class ScheduleLine:
    """IterationScheduleLine (TensorNetworks/IterationSchedule.H:13-38), variational fields."""
    max_iter: int
    Dmax: int
    DeltaD: int = 1
    eps_energy: float = 1e-12      # itsDelatEnergy1Epsilon
    eps_lambda: float = 1e-12      # itsDeltaLambdaEpsilon
    eps_eigen: float = 1e-12       # itsEigenSolverEpsilon (Lanczos residual)


@dataclass
class GSResult:
    de2: float = 0.0
    energy: float = 0.0
    nsweeps: int = 0
    history: list = field(default_factory=list)
    stats: SweepStats = None


def create_h2(Ws):
    """Hamiltonian::CreateH2Operator (TensorNetworks/Hamiltonian.C:8-16): H*H,
    CanonicalForm and Compress(Parker, 0, 1e-13) when H has a regular form;
    the plain product otherwise."""
    H2 = mpo_square(Ws)
    form = regular_form(Ws)
    if form is None:
        return H2
    return mpo_reduce(H2, form, ops=3, Dmax=0, eps=1e-13)[0]


def FindVariationalGroundState(chain, schedule, opts=None, H2=None, log=None):
    """Run the schedule on a Chain; the chain's own MPO is the Hamiltonian."""
    opts = opts or SweepOpts.default()
    H2 = H2 if H2 is not None else create_h2(chain.Ws)
    stats = SweepStats()
    res = GSResult(stats=stats)
    chain.normalize(1)
    chain.load_caches(opts)
    for line in schedule:
        o = SweepOpts.default()
        for f, _ in SweepOpts._fields_:
            setattr(o, f, getattr(opts, f))
        o.eig_tol = min(opts.eig_tol, line.eps_eigen) if line.eps_eigen else opts.eig_tol
        D1, D2 = chain.dims()
        Dstart = max(D2)
        D = Dstart
        while D <= line.Dmax:
            if D > Dstart:
                chain.increase_bond_dims(D)
                chain.normalize(1)
                chain.load_caches(o)
            it = 0
            while it < line.max_iter:
                chain.sweep(0, o, stats)
                chain.sweep(1, o, stats)
                dE, dl = chain.deltas()
                res.history.append((D, it, dE, dl, stats.energy))
                if log:
                    log(f"{it:4d} {D:4d} {dE:.1e} {dl:.1e}")
                if dE < line.eps_energy and dl < line.eps_lambda:
                    break
                it += 1
            it += 1
            res.nsweeps = it
            E1 = chain.expectation()
            E2 = chain.expectation(H2)
            res.energy = E1
            res.de2 = E2 - E1 * E1
            if log:
                log(f"Variational GS D={D:4d}, {it:4d} iterations, <E>={E1 / max(chain.L - 1, 1):.11f}, "
                    f"<E^2>-<E>^2={res.de2:.2e}")
            D += line.DeltaD
    return res
