"""CPU checks of the C ABI library: it loads, exports every entry point that
include/tn_abi.h declares, and its host-only logic matches the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import tensornetworks_amd as tn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tn_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tn_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("tn_heff_apply", "tn_env_update", "tn_eig_lowest", "tn_svd", "tn_chain_sweep",
              "tn_chain_sweep2", "tn_svcompress_rank"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tn.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_cpu_fallback_without_device():
    # Without a device the context refuses to open; nothing runs on the CPU.
    if tn.lib.tn_device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(tn.TnError):
        tn.Context(0)


@pytest.mark.parametrize("seed", range(6))
def test_svcompress_rule_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    s = np.sort(rng.uniform(0, 1, 20) ** 4)[::-1].copy()
    for Dmax, eps in ((5, 0.0), (0, 1e-3), (8, 1e-3), (30, 1e-8), (3, 0.5)):
        a = tn.svcompress_rank(s, Dmax, eps)
        b = O.svcompress(s, Dmax, eps)
        assert a[0] == b[0] and abs(a[1] - b[1]) < 1e-15 and abs(a[2] - b[2]) < 1e-12


def test_sweep_opts_defaults():
    o = tn.SweepOpts.default()
    assert o.precision == tn.TN_FP64 and o.max_krylov >= 16 and o.eig_tol > 0
    assert o.reorth == 0 and o.skip_identity == 1
    o = tn.SweepOpts.default(reorth=1, precond=3)
    assert o.reorth == 1 and o.precond == 3
    with pytest.raises(TypeError):
        tn.SweepOpts.default(reorht=1)
