"""Benchmark: DMRG sweep time and Heff TFLOP/s at D=1024, mixed precision vs FP64.

Workload (one GPU): S=1 Heisenberg open chain, L=100, d=3, MPO Dw=5, J=1,
single-ion anisotropy Dz=0.1 on Sz^2, bond dimension D=1024, seeded random
initial MPS (U(-1,1), normalised).  One step = one sweep = a left pass plus a
right pass = 2L one-site updates (eigensolve + SVD gauge move + environment
update), everything resident in HBM.

Warmup: W sweeps from the random state in the headline precision bring the
MPS close to its fixed point; the state is saved.  Timed region: K sweeps in
the headline precision (default "bf16x3": every Heff operand split into three
bfloat16 pieces, the six leading cross products on the bfloat16 matrix cores
with FP32 accumulation, FP32 environment updates; Krylov vectors,
re-orthogonalisation, tridiagonal solve, energies and QR/SVD gauge moves in
FP64).  From the identical saved state the same K sweeps are then timed with
every local solve refined by FP64 Lanczos iterations, with FP32 contractions,
with BF16 contractions refined by FP32 ones, and in FP64; |dE|/L compares the FP64 energies <H> of
the final MPS of the headline and FP64 runs.  value = algorithmic Heff flops
(identity channels counted, as if contracted) / device time of the Heff
applications (HIP events) during the timed sweeps.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"bf16": 2500.0, "bf16x3": 2500.0, "fp32": 157.3, "fp64": 78.6}  # dense MI355X TFLOP/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--L", type=int, default=100)
    p.add_argument("--D", type=int, default=1024)
    p.add_argument("--S", type=float, default=1.0)
    p.add_argument("--Dz", type=float, default=0.1)
    p.add_argument("--precision", default="bf16x3", choices=["mixed", "bf16", "bf16x3", "fp32", "fp64"],
                   help="fp32: FP32 contractions (FP32 environments); bf16x3: Heff contractions with every "
                        "operand split into three bfloat16 pieces, six cross products on the bfloat16 matrix "
                        "cores with FP32 accumulation (FP32 environment updates); mixed: BF16 contractions "
                        "refined by FP32 ones; Krylov vectors, Ritz solve, gauge and energies are FP64 in every "
                        "mode")
    p.add_argument("--no-bf16-compare", action="store_true",
                   help="skip the extra timed sweep of the BF16-then-FP32 scheme")
    p.add_argument("--eig-tol", type=float, default=1e-8)
    p.add_argument("--max-krylov", type=int, default=24)
    p.add_argument("--max-restarts", type=int, default=0)
    p.add_argument("--gauge", default="qr", choices=["qr", "svd"])
    p.add_argument("--reorth", default="partial", choices=["partial", "full"],
                   help="Lanczos re-orthogonalisation: partial (Simon's omega recurrence, a pass against the "
                        "whole basis only when needed) or full (every iteration); the FP64 comparison sweep "
                        "uses the same mode")
    p.add_argument("--no-skip-identity", action="store_true",
                   help="contract identity environment channels like the others instead of copying them")
    p.add_argument("--no-fp64-compare", action="store_true")
    p.add_argument("--no-fp32-compare", action="store_true",
                   help="skip the extra timed sweep with FP32 contractions (when the headline is another scheme)")
    p.add_argument("--no-refine-compare", action="store_true",
                   help="skip the extra timed sweep in the headline precision with every local solve refined "
                        "by FP64 Lanczos iterations")
    p.add_argument("--no-autotune", action="store_true", help="built-in GEMM tile choice instead of per-shape tuning")
    p.add_argument("--tune-db", default=os.path.join(ROOT, "tensornetworks_amd", "tuning", "mi355x_gemm.txt"),
                   help="saved GEMM tuning table to start from (shapes not in it are tuned on first use)")
    p.add_argument("--tune-save", default=None, help="write the final tuning table here")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--seed", type=int, default=20261202)
    p.add_argument("--out", default=None)
    p.add_argument("--workload", default="chain", choices=["chain", "batch", "precond", "twosite"],
                   help="chain: the D=1024 single-chain sweep (default); batch: 128 XXZ chains swept in lockstep; "
                        "precond: train the learned Davidson preconditioner on the device and compare local "
                        "solvers on the S=1/2 Heisenberg chain L=128, D=1024; twosite: two-site DMRG of the "
                        "S=1/2 Heisenberg chain L=64 with D ramped 32 -> 256, FP64, tolerance 1e-12")
    p.add_argument("--chains", type=int, default=128, help="batch workload: number of chains")
    p.add_argument("--no-sequential", action="store_true",
                   help="batch workload: skip the chain-by-chain comparison (for profiling runs)")
    p.add_argument("--precond-out", default=None, help="precond workload: write the trained network (.npz) here")
    p.add_argument("--heff-only", action="store_true",
                   help="set up the chain and time only the bulk-site Heff stages (for counter runs), "
                        "plus one plain device copy of known size that calibrates the byte counters")
    return p.parse_args()


def _median_s(fn, reps=3):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(D, d, L, matvecs_per_sweep, Dz):
    """Host-core baselines on bounded samples (medians of 3 runs each).

    MF-CPU (the value): the oracle's GEMM-ordered FP64 Heff apply (OpenMP) at
    the workload's bulk shape, in the unit of the GPU value; with the oracle's
    environment update at the same shape it gives a sweep-time estimate,
    matvecs-per-sweep x apply + 2L x environment update (gauge moves left
    out, so a lower bound).  RF-CPU: the reference's own local-solve scheme
    (GetHeff dense, then the lowest two eigenpairs with LAPACK, as
    LapackEigenSolver does with zheevx; here dsyevr on the real matrix) at
    the S=1/2 bulk shape of the reference's test chains, D in {8,16,32,48}."""
    import numpy as np
    import scipy.linalg as sla
    import oracle as O
    rng = np.random.default_rng(1)
    Dw = 5
    W = np.ascontiguousarray(O.xxz_bulk_W(d, 1.0, 1.0, 0.0, Dz, 0))
    Lv = rng.uniform(-1, 1, D * Dw * D)
    R = rng.uniform(-1, 1, Dw * D * D)
    x = rng.uniform(-1, 1, D * d * D)
    work = np.zeros(O.lib.orc_heff_work_size(d, D, D, Dw, Dw))
    y = np.zeros_like(x)
    envo = np.zeros(D * Dw * D)
    t_apply = _median_s(lambda: O.lib.orc_heff_apply_int(d, D, D, Dw, Dw, O.P(W), O.P(Lv), O.P(R), O.P(x), O.P(y),
                                                         O.P(work)))
    t_env = _median_s(lambda: O.lib.orc_env_left_int(d, D, D, Dw, Dw, O.P(W), O.P(Lv), O.P(x), O.P(envo),
                                                     O.P(work)))
    flops = 4.0 * d * Dw * D ** 3 + 2.0 * d * d * Dw * Dw * D * D
    rf = {}
    W2 = np.ascontiguousarray(O.xxz_bulk_W(2, 1.0, 1.0, 0.0, 0.0, 0))
    for Dr in (8, 16, 32, 48):
        Le = rng.uniform(-1, 1, Dr * Dw * Dr)
        Re = rng.uniform(-1, 1, Dw * Dr * Dr)

        def site():
            H = O.heff_dense(W2, Le, Re, 2, Dr, Dr)
            H = 0.5 * (H + H.T)  # random environments: symmetrised for the symmetric solver
            sla.eigh(H, subset_by_index=[0, 1], driver="evr", overwrite_a=True, check_finite=False)
        rf[str(Dr)] = _median_s(site)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": flops / t_apply / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": f"median of 3 FP64 Heff applies at D={D}, d={d}, Dw=5 (oracle GEMM-ordered, OpenMP), "
                      f"{t_apply:.3f} s each; environment update {t_env:.3f} s",
            "mf_cpu": {"heff_apply_s": t_apply, "env_update_s": t_env, "matvecs_per_sweep": matvecs_per_sweep,
                       "sweep_estimate_s": matvecs_per_sweep * t_apply + 2 * L * t_env,
                       "note": "matvecs of the FP64 GPU sweep x CPU apply + 2L x CPU environment update; "
                               "gauge moves not counted (lower bound)"},
            "rf_cpu": {"site_solve_s": rf,
                       "config1_sweep_estimate_s": 2 * 10 * rf["8"],
                       "note": "per bulk-site update of the S=1/2 chain: dense GetHeff (oracle) + lowest two "
                               "eigenpairs (LAPACK dsyevr, the real counterpart of zheevx IL=1 IU=2); the config-1 "
                               "estimate is 2L x the D=8 bulk site at L=10"}}


def run_batch(a, tn, ctx):
    """Batched workload: chains with L~U{16..128}, D~U{32..256}, J~U(0.5,1.5),
    Jz~U(0,2), hz~U(-0.5,0.5) (S=1/2 XXZ, open), FP32 contractions, FP64
    Krylov/Ritz/gauge/energy, QR gauge.  One step = one sweep (left and right
    pass) of every chain, all chains advanced in lockstep with grouped
    launches.  The same sweep is then timed chain by chain for comparison."""
    from tensornetworks_amd.models import xxz_mpo, batched_workload
    o = tn.SweepOpts.default(precision=tn.TN_FP32, env_precision=tn.TN_FP32, refine=0, gauge=1,
                             eig_tol=a.eig_tol, max_krylov=a.max_krylov, max_restarts=a.max_restarts,
                             reorth=1 if a.reorth == "partial" else 0)
    w = batched_workload(a.chains, a.seed)
    t0 = time.perf_counter()
    chains = []
    for k, (L, D, J, Jz, hz) in enumerate(w):
        ch = tn.Chain(ctx, xxz_mpo(L, 0.5, Jxy=J, Jz=Jz, hz=hz), D)
        ch.initialize(1, a.seed + 1 + k)
        ch.normalize_qr(1)
        ch.load_caches(o)
        chains.append(ch)
    batch = tn.Batch(ctx, chains, max_krylov=a.max_krylov)
    setup_s = time.perf_counter() - t0
    print(f"# batch setup {setup_s:.1f}s ({len(chains)} chains)", file=sys.stderr, flush=True)
    for wi in range(a.warmup):
        t1 = time.perf_counter()
        batch.sweep(0, o)
        batch.sweep(1, o)
        print(f"# warmup batched sweep {wi} {time.perf_counter() - t1:.2f}s", file=sys.stderr, flush=True)
    saved = [[(ch.dims()[0][i], ch.dims()[1][i], ch.site(i)) for i in range(ch.L)] for ch in chains]
    st = tn.SweepStats()
    t1 = time.perf_counter()
    for k in range(a.steps):
        batch.sweep(0, o, st)
        batch.sweep(1, o, st)
    wall = time.perf_counter() - t1
    E_b = [ch.expectation() for ch in chains]
    print(f"# timed {a.steps} batched sweep(s) {wall:.2f}s matvecs={st.n_matvec} full-basis passes="
          f"{st.n_full_reorth} qr={st.n_qr} fallback={st.n_qr_fallback}", file=sys.stderr, flush=True)
    if a.no_sequential:
        print(json.dumps({"metric": "batched DMRG sweep time (profiling run, no chain-by-chain comparison)",
                          "ms_per_step": wall / a.steps * 1e3}), flush=True)
        return
    # the same sweep(s) chain by chain from the saved states
    for ch, sv in zip(chains, saved):
        for i, (d1, d2, M) in enumerate(sv):
            ch.set_site(i, d1, d2, M)
        ch.load_caches(o)
    sts = tn.SweepStats()
    t1 = time.perf_counter()
    for ch in chains:
        for k in range(a.steps):
            ch.sweep(0, o, sts)
            ch.sweep(1, o, sts)
    wall_seq = time.perf_counter() - t1
    E_s = [ch.expectation() for ch in chains]
    dE = max(abs(x - y) / ch.L for x, y, ch in zip(E_b, E_s, chains))
    print(f"# chain-by-chain {wall_seq:.2f}s, max |dE|/L batched vs chain-by-chain {dE:.2e}", file=sys.stderr,
          flush=True)
    Ls = [x[0] for x in w]
    Ds = [x[1] for x in w]
    out = {
        "metric": "batched DMRG sweep time, 128 S=1/2 XXZ chains (L 16..128, D 32..256)",
        "value": len(chains) * a.steps / wall,
        "unit": "chain-sweeps/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded couplings, lengths, bond dimensions, random initial MPS)",
        "config": {"workload": f"{len(chains)} open S=1/2 XXZ chains, L~U{{16..128}}, D~U{{32..256}}, J~U(0.5,1.5), "
                               "Jz~U(0,2), hz~U(-0.5,0.5); one-site sweep, FP32 contractions, FP64 "
                               "Krylov/Ritz/QR gauge/energy, all chains in lockstep with grouped launches",
                   "chains": len(chains), "L_total": int(sum(Ls)), "L_max": max(Ls), "D_max": max(Ds),
                   "eig_tol": a.eig_tol, "max_krylov": a.max_krylov, "seed": a.seed, "reorth": a.reorth},
        "heff_tflops": st.heff_flops / max(st.heff_ms, 1e-12) / 1e9,
        "breakdown_ms": {"heff": st.heff_ms, "lanczos": st.lanczos_ms, "gauge": st.svd_ms, "env": st.env_ms,
                         "wall": st.wall_ms},
        "n_matvec": st.n_matvec, "n_full_reorth": st.n_full_reorth, "n_qr": st.n_qr,
        "n_qr_fallback": st.n_qr_fallback,
        "chain_by_chain": {"sweep_s": wall_seq / a.steps, "n_matvec": sts.n_matvec,
                           "speedup_of_batching": wall_seq / wall, "max_dE_per_site": dE},
        "setup_s": setup_s,
    }
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


def run_precond(a, tn, ctx):
    """Learned preconditioner: samples from random XXZ chains at D <= 128
    (couplings as the batched workload), device training (own forward /
    backward / Adam kernels), then one sweep of the S=1/2 Heisenberg chain
    (L=128, D=1024 by default; --L/--D override) from the same warm state with
    each local solver: Lanczos, Davidson without / with the diagonal / with
    the learned correction at every step / with the learned correction on the
    first step only.  Reported: matvecs (Krylov iterations), the solves that
    fell back to Lanczos, sweep time and |dE|/L against the Lanczos sweep."""
    from tensornetworks_amd import precond
    from tensornetworks_amd.models import xxz_mpo
    t0 = time.perf_counter()
    X, av, tv = precond.generate_samples(ctx, n_chains=24, seed=a.seed % 100000, L_range=(12, 32), D_range=(16, 128),
                                         iters=1, nsamp=1024, sites_per_chain=12)
    params, norm, (c0, c1) = precond.train(ctx, X, av, tv, 1024, rounds=10, steps=400, batch=4096, lr=2e-3)
    train_s = time.perf_counter() - t0
    ctx.set_mlp(params, norm)
    print(f"# trained on {len(av)} samples in {train_s:.1f}s: mean first-step cosine {c0:.3f} -> {c1:.3f}",
          file=sys.stderr, flush=True)
    L, D = (a.L if a.L != 100 else 128), a.D
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5), D)
    ch.initialize(1, a.seed)
    ch.normalize_qr(1)
    # FP64 contractions: Davidson's explicit residual |W y - theta u| has a floor
    # near eps * |theta| set by the operator's precision, which FP32 puts above
    # the tolerances of interest (Lanczos's implicit estimate has no such floor)
    base = dict(precision=tn.TN_FP64, env_precision=tn.TN_FP64, refine=0, gauge=1)
    ow = tn.SweepOpts.default(eig_tol=1e-6, max_krylov=24, max_restarts=0, **base)
    ch.load_caches(ow)
    for w in range(max(a.warmup, 1)):
        ch.sweep(0, ow)
        ch.sweep(1, ow)
    saved = [(ch.dims()[0][i], ch.dims()[1][i], ch.site(i)) for i in range(L)]
    res = {}
    for name, method, kind in [("lanczos", 0, 0), ("davidson_none", 1, 0), ("davidson_diag", 1, 1),
                               ("davidson_learned", 1, 2), ("davidson_learned_first", 1, 3)]:
        o = tn.SweepOpts.default(eig_tol=a.eig_tol, max_krylov=a.max_krylov, max_restarts=3, eig_method=method,
                                 precond=kind, **base)
        for i, (d1, d2, M) in enumerate(saved):
            ch.set_site(i, d1, d2, M)
        ch.load_caches(o)
        st = tn.SweepStats()
        t1 = time.perf_counter()
        ch.sweep(0, o, st)
        ch.sweep(1, o, st)
        res[name] = {"n_matvec": st.n_matvec, "sweep_s": time.perf_counter() - t1, "E_per_site": ch.expectation() / L,
                     "n_lanczos_fallback": st.n_precond_fallback}
        print(f"# {name}: {res[name]}", file=sys.stderr, flush=True)
    e_ref = res["lanczos"]["E_per_site"]
    for r in res.values():
        r["dE_per_site_vs_lanczos"] = abs(r["E_per_site"] - e_ref)
    out = {
        "metric": "Krylov matvecs per sweep and |dE|/L of the learned Davidson preconditioner vs plain Lanczos",
        "value": res["lanczos"]["n_matvec"] / max(res["davidson_learned_first"]["n_matvec"], 1),
        "unit": "x fewer matvecs than Lanczos (learned first correction, diagonal after)",
        "n_gpus": 1, "steps": 1, "warmup": max(a.warmup, 1),
        "ms_per_step": res["davidson_learned_first"]["sweep_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": "synthetic (training chains from a seed; evaluation from a seeded random MPS)",
        "config": {"workload": f"S=1/2 Heisenberg open chain L={L}, D={D}, one sweep per local solver from the same "
                               "warm state; FP64 contractions, Krylov/Ritz/QR gauge/energy",
                   "eig_tol": a.eig_tol, "max_krylov": a.max_krylov, "training_samples": int(len(av)),
                   "training_s": train_s, "first_step_cosine": [c0, c1]},
        "solvers": res,
    }
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    if a.precond_out:
        precond.save(a.precond_out, params, norm)


def run_twosite(a, tn, ctx):
    """Two-site DMRG: open S=1/2 Heisenberg chain L=64 (J=Jz=1, hz=0), seeded
    random MPS at D=32, two-site sweeps (left and right pass) at Dmax = 32, 64,
    128, 256, 256 with FP64 contractions and Lanczos tolerance 1e-12, no
    singular-value cutoff.  One step = one two-site sweep at D=256 (the last
    stage).  Then one-site sweeps at D=256 from the final state: their energy
    may only fall below the two-site one by the discarded weight's effect."""
    from tensornetworks_amd.models import xxz_mpo
    L = 64
    o = tn.SweepOpts.default(precision=tn.TN_FP64, env_precision=tn.TN_FP64, refine=0, eig_tol=1e-12,
                             max_krylov=a.max_krylov, max_restarts=8, gauge=0,
                             reorth=1 if a.reorth == "partial" else 0)
    ch = tn.Chain(ctx, xxz_mpo(L, 0.5), 32, Dcap=256)
    ch.initialize(1, a.seed)
    ch.normalize(1)
    ch.load_caches(o)
    stages = []
    last = None
    for Dmax in (32, 64, 128, 256, 256):
        st = tn.SweepStats()
        t1 = time.perf_counter()
        _, tr0 = ch.sweep2(0, Dmax, 0.0, o, st)
        _, tr1 = ch.sweep2(1, Dmax, 0.0, o, st)
        dt = time.perf_counter() - t1
        e = ch.expectation() / (L - 1)
        stages.append({"Dmax": Dmax, "sweep_s": dt, "E_per_bond": e, "truncation": max(tr0, tr1),
                       "n_matvec": st.n_matvec, "heff_tflops": st.heff_flops / max(st.heff_ms, 1e-12) / 1e9,
                       "lanczos_ms": st.lanczos_ms, "svd_ms": st.svd_ms})
        print(f"# two-site Dmax={Dmax}: {dt:.2f}s E/bond={e:.12f} truncation={max(tr0, tr1):.2e} "
              f"matvecs={st.n_matvec} dims max {max(ch.dims()[1])}", file=sys.stderr, flush=True)
        last = (dt, st)
    e2 = ch.expectation() / (L - 1)
    o1 = tn.SweepOpts.from_buffer_copy(o)
    ch.load_caches(o1)
    t1 = time.perf_counter()
    for _ in range(2):
        ch.sweep(0, o1)
        ch.sweep(1, o1)
    t_one = (time.perf_counter() - t1) / 2
    e1 = ch.expectation() / (L - 1)
    print(f"# one-site at D=256 from the two-site state: {t_one:.2f}s/sweep E/bond={e1:.12f} "
          f"(two-site {e2:.12f})", file=sys.stderr, flush=True)
    dt, st = last
    out = {
        "metric": "two-site DMRG sweep time at D=256, S=1/2 Heisenberg L=64 (D ramped 32 -> 256)",
        "value": 1.0 / dt,
        "unit": "sweeps/s",
        "n_gpus": 1, "steps": 1, "warmup": 4, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": "synthetic (seeded random initial MPS)",
        "config": {"workload": "open S=1/2 Heisenberg chain L=64, J=Jz=1, d=2, Dw=5, two-site sweeps at Dmax 32, "
                               "64, 128, 256, 256, FP64 contractions, Lanczos tolerance 1e-12, no cutoff",
                   "L": L, "eig_tol": 1e-12, "max_krylov": a.max_krylov, "reorth": a.reorth, "seed": a.seed},
        "stages": stages,
        "E_per_bond_two_site": e2,
        "E_per_bond_one_site_after": e1,
        "one_site_sweep_s": t_one,
        "heff_tflops": st.heff_flops / max(st.heff_ms, 1e-12) / 1e9,
    }
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; sharding across GPUs is designed "
                                   "but not launched", "n_gpus": a.gpus}))
        sys.exit(2)
    import tensornetworks_amd as tn
    from tensornetworks_amd.models import xxz_mpo

    P = {"bf16": tn.TN_BF16, "bf16x3": tn.TN_BF16X3, "fp32": tn.TN_FP32, "fp64": tn.TN_FP64}
    prec = tn.TN_BF16 if a.precision == "mixed" else P[a.precision]
    if a.workload in ("batch", "precond", "twosite"):
        ctx = tn.Context(0)
        if not a.no_autotune:
            ctx.set_autotune(True)
            if a.tune_db and os.path.exists(a.tune_db):
                ctx.tune_load(a.tune_db)
        {"batch": run_batch, "precond": run_precond, "twosite": run_twosite}[a.workload](a, tn, ctx)
        if a.tune_save and not a.no_autotune:
            ctx.tune_save(a.tune_save)
        return
    d = int(round(2 * a.S + 1))
    ctx = tn.Context(0)
    if not a.no_autotune:
        ctx.set_autotune(True)
        if a.tune_db and os.path.exists(a.tune_db):
            n = ctx.tune_load(a.tune_db)
            print(f"# loaded {n} tuned GEMM shapes from {os.path.relpath(a.tune_db, ROOT)}", file=sys.stderr,
                  flush=True)
    Ws = xxz_mpo(a.L, a.S, Dz=a.Dz)
    ch = tn.Chain(ctx, Ws, a.D)
    ch.initialize(1, a.seed)
    t0 = time.perf_counter()
    gauge = 1 if a.gauge == "qr" else 0
    if gauge:
        ch.normalize_qr(1)
    else:
        ch.normalize(1)
    common = dict(eig_tol=a.eig_tol, max_krylov=a.max_krylov, max_restarts=a.max_restarts, gauge=gauge,
                  skip_identity=0 if a.no_skip_identity else 1, reorth=1 if a.reorth == "partial" else 0)
    if a.precision == "mixed":
        o_mixed = tn.SweepOpts.default(precision=tn.TN_BF16, env_precision=tn.TN_FP32, refine=1,
                                       refine_precision=tn.TN_FP32, refine_tol=a.eig_tol, **common)
    elif a.precision == "bf16x3":
        o_mixed = tn.SweepOpts.default(precision=prec, env_precision=tn.TN_FP32, refine=0, **common)
    else:
        o_mixed = tn.SweepOpts.default(precision=prec, env_precision=prec, refine=0, **common)
    o_fp64 = tn.SweepOpts.default(precision=tn.TN_FP64, env_precision=tn.TN_FP64, refine=0, **common)
    # the headline scheme with the Ritz vector of every local solve refined by
    # FP64 Lanczos iterations (FP64 contractions) to eig_tol
    o_ref64 = tn.SweepOpts.from_buffer_copy(o_mixed)
    o_ref64.refine = 1
    o_ref64.refine_precision = tn.TN_FP64
    o_ref64.refine_tol = a.eig_tol
    o_bf16 = tn.SweepOpts.default(precision=tn.TN_BF16, env_precision=tn.TN_FP32, refine=1,
                                  refine_precision=tn.TN_FP32, refine_tol=a.eig_tol, **common)
    ch.load_caches(o_mixed)
    setup_s = time.perf_counter() - t0
    print(f"# setup {setup_s:.1f}s", file=sys.stderr, flush=True)
    if a.heff_only:
        site = a.L // 2
        s1, mix, s3, f1, f3 = ch.heff_bench(site, prec, reps=10, skip_identity=not a.no_skip_identity)
        print(f"# heff bench site {site} dims {ch.dims()[0][site]}x{ch.dims()[1][site]}: stage1 {s1:.4f} ms "
              f"stage3 {s3:.4f} ms mix {mix:.4f} ms", file=sys.stderr, flush=True)
        if not a.no_autotune:
            print("# tuned GEMM shapes:\n" + ctx.tune_report(), file=sys.stderr, flush=True)
        n = 64 << 20  # 64 Mi doubles = 512 MiB read + 512 MiB written
        ms = ctx.copy_bench(n)
        print(f"# calibration copy (scale_copy_k): {n * 8} bytes read, {n * 8} bytes written, {ms:.3f} ms "
              f"({2 * n * 8 / ms / 1e6:.0f} GB/s)", file=sys.stderr, flush=True)
        return
    for w in range(a.warmup):
        ch.sweep(0, o_mixed)
        ch.sweep(1, o_mixed)
        print(f"# warmup sweep {w} E/L={ch.expectation() / a.L:.12f} ({time.perf_counter() - t0:.1f}s)",
              file=sys.stderr, flush=True)
    D1, D2 = ch.dims()
    saved = [(D1[i], D2[i], ch.site(i)) for i in range(a.L)]
    # the timed sweeps record no events: the Heff device time comes from a
    # second, instrumented run of the same sweeps from the same state
    o_timed = tn.SweepOpts.from_buffer_copy(o_mixed)
    o_timed.time_heff = 0
    # ---- timed region: K sweeps in the headline precision ----
    st = tn.SweepStats()
    t0 = time.perf_counter()
    for k in range(a.steps):
        ch.sweep(0, o_timed, st)
        ch.sweep(1, o_timed, st)
    wall = time.perf_counter() - t0
    # ---------------------------------------------------------
    E_mixed = ch.expectation()
    print(f"# timed {a.steps} sweep(s) {wall:.2f}s E/L={E_mixed / a.L:.12f} matvecs={st.n_matvec} "
          f"(fp64 {st.n_matvec_fp64}) full-basis passes={st.n_full_reorth} qr={st.n_qr} "
          f"shifted={st.n_qr_shifted} fallback={st.n_qr_fallback} local solves converged "
          f"{st.n_solve_converged} capped {st.n_solve_capped}", file=sys.stderr, flush=True)

    def timed_from_saved(opts, tune=True):
        """K sweeps from the saved state (after one untimed sweep that tunes
        the GEMM shapes of this precision when autotuning is on)."""
        def restore():
            for i, (d1, d2, M) in enumerate(saved):
                ch.set_site(i, d1, d2, M)
            ch.load_caches(opts)
        restore()
        if tune and not a.no_autotune:
            ch.sweep(0, opts)
            ch.sweep(1, opts)
            restore()
        stx = tn.SweepStats()
        t1 = time.perf_counter()
        for k in range(a.steps):
            ch.sweep(0, opts, stx)
            ch.sweep(1, opts, stx)
        return stx, time.perf_counter() - t1, ch.expectation()

    # instrumented rerun of the timed sweeps: HIP events around every Heff apply
    sti, walli, _ = timed_from_saved(o_mixed, tune=False)
    print(f"# instrumented rerun {walli:.2f}s matvecs={sti.n_matvec} heff {sti.heff_ms:.0f} ms "
          f"({sti.heff_flops / max(sti.heff_ms, 1e-12) / 1e9:.1f} TFLOP/s)", file=sys.stderr, flush=True)
    site = a.L // 2
    s1, mix, s3, f1, f3 = ch.heff_bench(site, prec, reps=10, skip_identity=not a.no_skip_identity)
    print(f"# heff bench site {site}: stage1 {s1:.3f} ms stage3 {s3:.3f} ms mix {mix:.3f} ms",
          file=sys.stderr, flush=True)
    res_fp64 = None
    res_bf16 = None
    res_ref64 = None
    res_fp32 = None
    if a.precision not in ("fp32", "fp64") and not a.no_fp32_compare:
        o_fp32 = tn.SweepOpts.default(precision=tn.TN_FP32, env_precision=tn.TN_FP32, refine=0, **common)
        st32, wall32, E32 = timed_from_saved(o_fp32)
        res_fp32 = {"scheme": "FP32 contractions, FP32 environments", "sweep_s": wall32 / a.steps,
                    "E_per_site": E32 / a.L, "n_matvec": st32.n_matvec,
                    "heff_tflops": st32.heff_flops / max(st32.heff_ms, 1e-12) / 1e9}
        print(f"# fp32 {wall32:.2f}s E/L={E32 / a.L:.12f} matvecs={st32.n_matvec}", file=sys.stderr, flush=True)
    if a.precision != "fp64" and not a.no_refine_compare:
        str_, wallr, Er = timed_from_saved(o_ref64)
        res_ref64 = {"scheme": f"{a.precision} contractions, every local solve refined by FP64 Lanczos to eig_tol",
                     "sweep_s": wallr / a.steps, "E_per_site": Er / a.L, "n_matvec": str_.n_matvec,
                     "n_matvec_fp64": str_.n_matvec_fp64, "n_solve_converged": str_.n_solve_converged,
                     "n_solve_capped": str_.n_solve_capped}
        print(f"# {a.precision} + fp64 refine {wallr:.2f}s E/L={Er / a.L:.12f} matvecs={str_.n_matvec} "
              f"(fp64 {str_.n_matvec_fp64})", file=sys.stderr, flush=True)
    if a.precision != "mixed" and not a.no_bf16_compare:
        stb, wallb, Eb = timed_from_saved(o_bf16)
        res_bf16 = {"scheme": "BF16 contractions, FP32 environments, FP32-contraction refinement",
                    "sweep_s": wallb / a.steps, "E_per_site": Eb / a.L, "n_matvec": stb.n_matvec,
                    "heff_tflops": stb.heff_flops / max(stb.heff_ms, 1e-12) / 1e9}
        print(f"# bf16+fp32 refine {wallb:.2f}s E/L={Eb / a.L:.12f} matvecs={stb.n_matvec}", file=sys.stderr,
              flush=True)
    if not a.no_fp64_compare:
        st64, wall64, E64 = timed_from_saved(o_fp64)
        res_fp64 = {"sweep_s": wall64 / a.steps, "E_per_site": E64 / a.L,
                    "heff_tflops": st64.heff_flops / max(st64.heff_ms, 1e-12) / 1e9,
                    "n_matvec": st64.n_matvec, "n_full_reorth": st64.n_full_reorth,
                    "dE_per_site": abs(E_mixed - E64) / a.L,
                    "gauge_ms": st64.svd_ms, "env_ms": st64.env_ms, "lanczos_ms": st64.lanczos_ms,
                    "heff_ms": st64.heff_ms, "n_qr_fallback": st64.n_qr_fallback,
                    "n_solve_converged": st64.n_solve_converged, "n_solve_capped": st64.n_solve_capped,
                    "ms_per_matvec": st64.heff_ms / max(st64.n_matvec, 1)}
        print(f"# fp64 {wall64:.2f}s E/L={E64 / a.L:.12f} |dE|/L={abs(E_mixed - E64) / a.L:.3e}",
              file=sys.stderr, flush=True)
    # the split product's stage 1 carries the MPO stage in its epilogue
    # (x3_stage1_mix_k, 96 x 128 tiles: 32 a' x 3 n rows, 4 kept channels x 32 b
    # columns) wherever the library covers the shape: its events then show no
    # separate MPO stage
    fused = a.precision == "bf16x3" and mix < 0.1 * s1
    stage1_tile = "fused 96x128" if fused else None
    if not a.no_autotune:
        report = ctx.tune_report()
        print("# tuned GEMM shapes:\n" + report, file=sys.stderr, flush=True)
    if not a.no_autotune and not fused:
        # tile of the bulk-site stage-1 GEMM: x (D*d x D) times R^T (nR*D x D),
        # nR = right channels contracted (Dw, or Dw - 1 with the identity copied)
        tname = {"fp32": "f32->f32", "fp64": "f64->f64", "bf16x3": "bf16x3->f32"}.get(a.precision, "bf16->f32")
        nR = int(round(f1 / (2.0 * d * a.D ** 3)))
        key = f"{tname} KK M={a.D * d} N={nR * a.D} K={a.D} "
        for line in report.splitlines():
            if line.startswith(key):
                stage1_tile = line.split(": tile ")[1]
    if not a.no_autotune and a.tune_save:
        ctx.tune_save(a.tune_save)
    # HBM bytes per launch of the stage-1 GEMM from the committed byte-counter
    # summary, used only when it describes the same kernel configuration
    traffic = None
    tp = os.path.join(ROOT, "profiles", "heff_stage1_traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            tr = json.load(f)
        if (tr.get("precision") == a.precision and tr.get("D") == a.D and tr.get("d") == d
                and tr.get("tile") == stage1_tile):
            traffic = {"bytes": tr["bytes_per_launch"], "source": "profiles/heff_stage1_traffic.json"}
    value = sti.heff_flops / max(sti.heff_ms, 1e-12) / 1e9
    # the split-bfloat16 scheme's arithmetic is its six bfloat16 products per
    # FP32-equivalent product: those are what the matrix cores must execute
    nprod = 6 if a.precision == "bf16x3" else 1
    achieved = nprod * f1 / (s1 * 1e-3) / 1e12
    hp = "bf16" if a.precision == "mixed" else a.precision
    B1 = {"bf16": 2, "bf16x3": 6, "fp32": 4, "fp64": 8}[hp]
    out = {
        "metric": "DMRG sweep time + Heff TFLOP/s at D=1024; mixed vs FP64 at |dE/L|<=1e-8",
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": hp,
        "data": "synthetic (seeded random initial MPS)",
        "config": {"workload": f"S={a.S:g} Heisenberg + Dz={a.Dz:g} Sz^2 open chain, L={a.L}, d={d}, Dw=5, D={a.D}, "
                               f"one-site sweep, precision {a.precision} (Heff {hp}"
                               + (", FP32 environments, FP32-contraction refinement" if a.precision == "mixed" else "")
                               + "), FP64 Krylov/Ritz/gauge/energy",
                   "precision": a.precision,
                   "L": a.L, "D": a.D, "d": d, "Dw": 5, "eig_tol": a.eig_tol, "max_krylov": a.max_krylov,
                   "max_restarts": a.max_restarts, "gauge": a.gauge, "autotune": not a.no_autotune,
                   "skip_identity": not a.no_skip_identity, "reorth": a.reorth},
        "sweep_s": wall / a.steps,
        "E_per_site": E_mixed / a.L,
        "breakdown_ms": {"lanczos": st.lanczos_ms, "env": st.env_ms, "gauge": st.svd_ms, "wall": st.wall_ms},
        "instrumented_rerun": {"wall_s": walli, "heff_ms": sti.heff_ms, "heff_gemm_ms": sti.heff_gemm_ms,
                               "lanczos_ms": sti.lanczos_ms, "n_matvec": sti.n_matvec,
                               "note": "the timed sweeps again from the same state with HIP events around the "
                                       "Heff stages; value and ms_per_matvec come from it"},
        "n_qr": st.n_qr, "n_qr_fallback": st.n_qr_fallback, "n_qr_shifted": st.n_qr_shifted,
        "n_matvec": st.n_matvec, "n_matvec_fp64": st.n_matvec_fp64, "n_full_reorth": st.n_full_reorth,
        "heff_gemm_tflops_launched": sti.heff_gemm_flops / max(sti.heff_gemm_ms, 1e-12) / 1e9,
        "n_ident_left": st.n_ident_left, "n_ident_right": st.n_ident_right,
        "n_solve_converged": st.n_solve_converged, "n_solve_capped": st.n_solve_capped,
        "ms_per_matvec": sti.heff_ms / max(sti.n_matvec, 1),
        "fp64": res_fp64,
        "bf16_then_fp32": res_bf16,
        "fp64_refined": res_ref64,
        "fp32": res_fp32,
        "speedup_vs_fp64": (res_fp64["sweep_s"] / (wall / a.steps)) if res_fp64 else None,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAKS[hp], "unit": "TFLOP/s",
                     "frac": achieved / PEAKS[hp], "traffic": traffic["bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None, "stage1_tile": stage1_tile,
                     "kernel": ("Heff stage-1 GEMM (psi x R, NT) with the MPO stage in its epilogue "
                                "(x3_stage1_mix_k), bulk site" if fused else
                                "Heff stage-1 GEMM (psi x R, NT) at the bulk site"),
                     "per_launch_flops": nprod * f1, "per_launch_flops_fp32_equivalent": f1,
                     "products_per_fp32_product": nprod, "avg_launch_ms": s1,
                     "stage3_ms": s3, "stage3_flops": f3, "mpo_stage_ms": mix,
                     "algorithmic_bytes_B1": B1 * (2 * 5 * a.D ** 2 + 2 * d * a.D ** 2) + 8 * d * d * 25},
        "setup_s": setup_s,
    }
    if not a.no_cpu_baseline:
        mv = (res_fp64["n_matvec"] if res_fp64 else st.n_matvec) / a.steps
        out["cpu_baseline"] = cpu_baseline(a.D, d, a.L, mv, a.Dz)
        out["cpu_baseline"]["gpu_sweep_speedup_vs_mf_cpu_estimate"] = (
            out["cpu_baseline"]["mf_cpu"]["sweep_estimate_s"] / (wall / a.steps))
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
