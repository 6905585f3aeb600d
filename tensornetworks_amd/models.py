"""MPO construction for the nearest-neighbour spin chains on the sweep path.

Spin operators follow TensorNetworksImp/SpinCalculator.C:20-70 (index n <->
Sz = n - S; S+(m, n) != 0 iff m == n + 1).  The operator-valued matrix W of the
XXZ chain follows TensorNetworksImp/Hamiltonians/Hamiltonian_1D_NN_Heisenberg.C:35-79,
with a single-ion term Dz*Sz^2 added to the on-site slot (an extension used by
the S=1 workload), and the edge rows/columns of
Operators/SiteOperatorImp.C:33-51,109-134 (Upper: left keeps row 0, right keeps
column Dw-1; Lower: left keeps row Dw-1, right keeps column 0).
"""
import numpy as np

REGULAR_UPPER, REGULAR_LOWER = 0, 1


def spin_ops(d):
    S = 0.5 * (d - 1)
    Sp = np.zeros((d, d))
    Sm = np.zeros((d, d))
    Sz = np.zeros((d, d))
    for m in range(d):
        for n in range(d):
            sm, sn = m - S, n - S
            if m == n + 1:
                Sp[m, n] = np.sqrt(S * (S + 1) - sm * sn)
            if m + 1 == n:
                Sm[m, n] = np.sqrt(S * (S + 1) - sm * sn)
            if m == n:
                Sz[m, n] = sm
    return Sp, Sm, Sz


def xxz_bulk_W(d, Jxy=1.0, Jz=1.0, hz=0.0, Dz=0.0, form=REGULAR_UPPER):
    Sp, Sm, Sz = spin_ops(d)
    I = np.eye(d)
    On = hz * Sz + Dz * Sz @ Sz
    W = np.zeros((5, 5, d, d))
    if form == REGULAR_UPPER:
        W[0, 0], W[0, 1], W[0, 2], W[0, 3], W[0, 4] = I, Sp, Sm, Sz, On
        W[1, 4], W[2, 4], W[3, 4], W[4, 4] = 0.5 * Jxy * Sm, 0.5 * Jxy * Sp, Jz * Sz, I
    else:
        W[0, 0], W[1, 0], W[2, 0], W[3, 0], W[4, 0] = I, Sp, Sm, Sz, On
        W[4, 1], W[4, 2], W[4, 3], W[4, 4] = 0.5 * Jxy * Sm, 0.5 * Jxy * Sp, Jz * Sz, I
    return W


def edge_trim(Wb, L, form=REGULAR_UPPER):
    Dw = Wb.shape[0]
    lindex = 0 if form == REGULAR_UPPER else Dw - 1
    rindex = Dw - 1 if form == REGULAR_UPPER else 0
    Ws = []
    for i in range(L):
        W = Wb
        if i == 0:
            W = W[lindex:lindex + 1]
        if i == L - 1:
            W = W[:, rindex:rindex + 1]
        Ws.append(np.ascontiguousarray(W))
    return Ws


def xxz_mpo(L, S=0.5, Jxy=1.0, Jz=1.0, hz=0.0, Dz=0.0, form=REGULAR_UPPER):
    """Per-site W[i] (Dw1, Dw2, d, d) of H = sum J(SxSx+SySy)+Jz SzSz + sum hz Sz + Dz Sz^2."""
    d = int(round(2 * S + 1))
    return edge_trim(xxz_bulk_W(d, Jxy, Jz, hz, Dz, form), L, form)


def mpo_square(Ws):
    """Uncompressed product MPO H*H (OperatorValuedMatrix.H:410 TensorProduct):
    W2[(w1,v1),(w2,v2)] = W(w1,w2) @ W(v1,v2) in physical space."""
    out = []
    for W in Ws:
        a, b, d, _ = W.shape
        W2 = np.einsum("ijmk,uvkn->iujvmn", W, W).reshape(a * a, b * b, d, d)
        out.append(np.ascontiguousarray(W2))
    return out


def identity_mpo(L, d):
    return [np.eye(d).reshape(1, 1, d, d).copy() for _ in range(L)]


def random_xxz_couplings(n, seed):
    """Batched workload couplings: J~U(0.5,1.5), Jz~U(0,2), hz~U(-0.5,0.5)."""
    rng = np.random.default_rng(seed)
    return [(rng.uniform(0.5, 1.5), rng.uniform(0.0, 2.0), rng.uniform(-0.5, 0.5)) for _ in range(n)]


def batched_workload(n, seed, L_range=(16, 128), D_range=(32, 256)):
    """Chains of the batched workload: per chain (L, D, J, Jz, hz) with
    L uniform in L_range, D uniform in D_range, J~U(0.5,1.5), Jz~U(0,2),
    hz~U(-0.5,0.5), all from one seed (S=1/2 XXZ, open ends)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        L = int(rng.integers(L_range[0], L_range[1] + 1))
        D = int(rng.integers(D_range[0], D_range[1] + 1))
        out.append((L, D, rng.uniform(0.5, 1.5), rng.uniform(0.0, 2.0), rng.uniform(-0.5, 0.5)))
    return out
