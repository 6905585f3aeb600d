"""tensornetworks_amd -- MI355X (gfx950) variational MPS ground-state sweep.

Python view of the C ABI in include/tn_abi.h (libtn_hip.so, built in-tree by
``python tensornetworks_amd/build.py``).  There is no CPU fallback: importing
the package without the built library raises.
"""
from ._abi import (  # noqa: F401
    LIB_PATH,
    TN_BF16,
    TN_BF16X3,
    TN_FP32,
    TN_FP64,
    MLP_NP,
    Batch,
    Chain,
    Context,
    SweepOpts,
    SweepStats,
    TnError,
    lib,
    mpo_reduce,
    read_checkpoint_mpo,
    svcompress_rank,
)
from .models import (  # noqa: F401
    xxz_mpo, tfim_mpo, long_range_mpo, three_body_mpo, mpo_square, identity_mpo, mpo_contract)
from .dmrg import FindVariationalGroundState, ScheduleLine  # noqa: F401

__all__ = [
    "Context", "Chain", "Batch", "SweepOpts", "SweepStats", "TnError", "lib", "LIB_PATH",
    "TN_FP64", "TN_FP32", "TN_BF16", "TN_BF16X3", "MLP_NP", "svcompress_rank", "mpo_reduce", "read_checkpoint_mpo",
    "xxz_mpo", "tfim_mpo", "long_range_mpo", "three_body_mpo", "mpo_square", "identity_mpo", "mpo_contract",
    "FindVariationalGroundState", "ScheduleLine",
]
