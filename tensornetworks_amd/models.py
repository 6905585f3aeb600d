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


def sx_op(d):
    """Sx = (S+ + S-)/2 (TensorNetworksImp/SpinCalculator.C:49-52)."""
    Sp, Sm, _ = spin_ops(d)
    return 0.5 * (Sp + Sm)


def tfim_bulk_W(d, J=1.0, hx=0.0, form=REGULAR_UPPER):
    """Transverse-field Ising chain H = J sum Sz Sz + hx sum Sx, Dw = 3
    (TensorNetworksImp/Hamiltonians/Hamiltonian_1D_NN_TransverseIsing.C:26-46)."""
    _, _, Sz = spin_ops(d)
    I, Sx = np.eye(d), sx_op(d)
    W = np.zeros((3, 3, d, d))
    if form == REGULAR_UPPER:
        W[0, 0], W[0, 1], W[0, 2], W[1, 2], W[2, 2] = I, Sz, hx * Sx, J * Sz, I
    else:
        W[0, 0], W[1, 0], W[2, 0], W[2, 1], W[2, 2] = I, Sz, hx * Sx, J * Sz, I
    return W


def long_range_bulk_W(d, J=1.0, hx=0.0, NN=1, three_body=False, form=REGULAR_UPPER):
    """Long-range chains with Dw = 2 + NN(NN+1)/2: for each range i = 1..NN a
    string of i channels carries the left operator i sites to the right.
    Two-body (Hamiltonian_1D_2Body_LongRange.C:26-63): H = sum_i J/i^2 Sz_k Sz_{k+i}
    + hx sum Sx.  Three-body (Hamiltonian_1D_3Body_LongRange.C:26-63): the string
    opens with Sz^2, its links after the first multiply by Sz/j^2, it closes with
    J/i^2 Sz^2; + hx sum Sx."""
    _, _, Sz = spin_ops(d)
    I, Sx = np.eye(d), sx_op(d)
    Dw = 2 + NN * (NN + 1) // 2
    W = np.zeros((Dw, Dw, d, d))
    W[0, 0] = I
    W[Dw - 1, Dw - 1] = I
    opn = Sz @ Sz if three_body else Sz
    ib = 0
    for i in range(1, NN + 1):
        for j in range(1, i):
            link = (Sz / (j * j)) if three_body else I
            if form == REGULAR_UPPER:
                W[ib + j, ib + 1 + j] = link
            else:
                W[ib + 1 + j, ib + j] = link
        if form == REGULAR_UPPER:
            W[0, ib + 1] = opn
            W[ib + i, Dw - 1] = J / (i * i) * opn
        else:
            W[ib + 1, 0] = opn
            W[Dw - 1, ib + i] = J / (i * i) * opn
        ib += i
    if form == REGULAR_UPPER:
        W[0, Dw - 1] = hx * Sx
    else:
        W[Dw - 1, 0] = hx * Sx
    return W


def three_body_bulk_W(d, J=1.0, K=1.0, hz=0.0, form=REGULAR_UPPER):
    """The three-body chain of TensorNetworksImp/Hamiltonians/Hamiltonian_1D_3Body.C:26-56,
    Dw = 5, entries as given there (channels 2 and 3 are only linked to each
    other, so only J sum Sx Sx + hz sum Sz acts; the compression drops them)."""
    _, _, Sz = spin_ops(d)
    I, Sx = np.eye(d), sx_op(d)
    W = np.zeros((5, 5, d, d))
    if form == REGULAR_UPPER:
        W[0, 0], W[0, 1], W[0, 2], W[0, 4] = I, Sx, Sx, hz * Sz
        W[3, 2], W[1, 4], W[3, 4], W[4, 4] = Sz, J * Sx, K * Sx, I
    else:
        W[0, 0], W[1, 0], W[2, 0], W[4, 0] = I, Sx, Sx, hz * Sz
        W[2, 3], W[4, 1], W[4, 3], W[4, 4] = Sz, J * Sx, K * Sx, I
    return W


def tfim_mpo(L, S=0.5, J=1.0, hx=0.0, form=REGULAR_UPPER):
    d = int(round(2 * S + 1))
    return edge_trim(tfim_bulk_W(d, J, hx, form), L, form)


def long_range_mpo(L, S=0.5, J=1.0, hx=0.0, NN=1, three_body=False, form=REGULAR_UPPER):
    d = int(round(2 * S + 1))
    return edge_trim(long_range_bulk_W(d, J, hx, NN, three_body, form), L, form)


def three_body_mpo(L, S=0.5, J=1.0, K=1.0, hz=0.0, form=REGULAR_UPPER):
    d = int(round(2 * S + 1))
    return edge_trim(three_body_bulk_W(d, J, K, hz, form), L, form)


def mpo_contract(Ws):
    """Dense operator of an MPO (d^L x d^L, site 1 the most significant
    index): the brute-force check of small MPOs."""
    T = Ws[0][0]  # (Dw2, d, d)
    for W in Ws[1:]:
        # T[w][M][N] x W[w][v][m][n] -> [v][(M,m)][(N,n)]
        a, b, d, _ = W.shape
        n = T.shape[1]
        T = np.einsum("wMN,wvmn->vMmNn", T, W).reshape(b, n * d, n * d)
    return T[0]


def regular_form(Ws):
    """REGULAR_UPPER or REGULAR_LOWER when the MPO has that regular form (its
    edge blocks hold the identity where the form puts it), else None."""
    first, last = Ws[0], Ws[-1]
    d = first.shape[2]
    I = np.eye(d)
    if len(Ws) < 2 or first.shape[1] < 2:
        return None
    if np.array_equal(first[0, 0], I) and np.array_equal(last[-1, 0], I):
        return REGULAR_UPPER
    if np.array_equal(first[0, -1], I) and np.array_equal(last[0, 0], I):
        return REGULAR_LOWER
    return None


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
