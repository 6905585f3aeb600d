"""Synthetic workload generators (host only)."""
import numpy as np

from tensornetworks_amd.models import batched_workload


def test_batch_workload_generator():
    w = batched_workload(128, 20261202)
    assert len(w) == 128
    Ls = np.array([x[0] for x in w])
    Ds = np.array([x[1] for x in w])
    assert Ls.min() >= 16 and Ls.max() <= 128 and Ds.min() >= 32 and Ds.max() <= 256
    J = np.array([x[2] for x in w])
    Jz = np.array([x[3] for x in w])
    hz = np.array([x[4] for x in w])
    assert (J >= 0.5).all() and (J <= 1.5).all() and (Jz >= 0).all() and (Jz <= 2).all()
    assert (np.abs(hz) <= 0.5).all()
    assert w == batched_workload(128, 20261202)
