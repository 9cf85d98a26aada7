"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

Independent of the oracle and of the product: plain Python loops and numpy /
scipy library routines only.  The reference's own tests hold no golden vectors
for this path (its examples print and never assert), so these are built from
definitions, library routines and published known-answer vectors:

* mttkrp_*.npz: brute-force MTTKRP by definition (M_n(i_n, r) = sum over all
  other indices of X * prod_k A_k(i_k, r)) on integer data, so every value is
  exact in FP64 (3x4x5 R=2 and 2x3x4x5 R=3, the sizes named for the path).
* unfold_*.npz: mode-n unfoldings of the 1..prod(I) fill used by the
  reference's customTensor (DataGeneration.hpp:761-820), by index definition
  X_(n)(i_n, j), j = the other indices in increasing mode order, lowest fastest
  (Matricization.hpp:101-124), and the partial Khatri-Rao product in the same
  row order (PartialKhatriRao.hpp:59-227).
* eig_extremes.npz: extreme eigenvalues of symmetric PSD R x R matrices from
  numpy.linalg.eigvalsh (LAPACK), incl. Hadamard products of Grams.
* nnls_prox.npz: scipy.optimize.nnls solutions of the per-row proximal
  problem min_a 1/2 a^T Zt a - wt^T a, a >= 0, that NesterovMNLS solves.
* philox_kat.json: Philox4x32-10 known-answer vectors published with
  Random123 (Salmon et al., SC'11).
"""
import itertools
import json
import os

import numpy as np
import scipy.linalg
import scipy.optimize

HERE = os.path.dirname(os.path.abspath(__file__))


def brute_mttkrp(X, A, mode):
    dims = X.shape
    R = A[0].shape[1]
    M = np.zeros((dims[mode], R))
    for idx in itertools.product(*[range(d) for d in dims]):
        for r in range(R):
            p = X[idx]
            for k, i in enumerate(idx):
                if k != mode:
                    p *= A[k][i, r]
            M[idx[mode], r] += p
    return M


def mttkrp_fixture(dims, R, seed):
    rng = np.random.default_rng(seed)
    X = rng.integers(-3, 4, size=dims).astype(np.float64)
    A = [rng.integers(0, 4, size=(I, R)).astype(np.float64) for I in dims]
    out = {"dims": np.array(dims), "X": X.ravel(order="F")}
    for k in range(len(dims)):
        out["A%d" % k] = A[k]
        out["M%d" % k] = brute_mttkrp(X, A, k)
    name = "mttkrp_%s_r%d.npz" % ("x".join(map(str, dims)), R)
    np.savez(os.path.join(HERE, name), **out)


def unfold_fixture(dims, R):
    n = int(np.prod(dims))
    lin = np.arange(1, n + 1, dtype=np.float64)          # customTensor fill 1..prod(I)
    out = {"dims": np.array(dims), "X": lin}
    rng = np.random.default_rng(7)
    A = [rng.integers(1, 5, size=(I, R)).astype(np.float64) for I in dims]
    for k in range(len(dims)):
        out["A%d" % k] = A[k]
    for mode in range(len(dims)):
        others = [k for k in range(len(dims)) if k != mode]
        J = n // dims[mode]
        U = np.zeros((dims[mode], J))
        KR = np.zeros((J, R))
        for idx in itertools.product(*[range(d) for d in dims]):
            lin_idx = 0
            stride = 1
            for k in range(len(dims)):
                lin_idx += idx[k] * stride
                stride *= dims[k]
            j = 0
            s = 1
            for k in others:
                j += idx[k] * s
                s *= dims[k]
            U[idx[mode], j] = lin[lin_idx]
            for r in range(R):
                p = 1.0
                for k in others:
                    p *= A[k][idx[k], r]
                KR[j, r] = p
        out["U%d" % mode] = U
        out["KR%d" % mode] = KR
    np.savez(os.path.join(HERE, "unfold_%s.npz" % "x".join(map(str, dims))), **out)


def eig_fixture():
    rng = np.random.default_rng(11)
    mats = []
    for R, kind in [(5, "gram"), (10, "hadamard"), (32, "hadamard"), (64, "hadamard"), (16, "illcond"), (3, "diag")]:
        if kind == "gram":
            B = rng.random((20, R))
            V = B.T @ B
        elif kind == "hadamard":
            G1 = rng.random((50, R)); G2 = rng.random((60, R))
            V = (G1.T @ G1) * (G2.T @ G2)
        elif kind == "illcond":
            Q, _ = np.linalg.qr(rng.standard_normal((R, R)))
            ev = np.logspace(0, -7, R)
            V = (Q * ev) @ Q.T
            V = 0.5 * (V + V.T)
        else:
            V = np.diag([3.0, 1e-3, 7.5])
        w = np.linalg.eigvalsh(V)
        mats.append((V, w))
    out = {}
    for i, (V, w) in enumerate(mats):
        out["V%d" % i] = V
        out["w%d" % i] = w
    out["count"] = np.array(len(mats))
    np.savez(os.path.join(HERE, "eig_extremes.npz"), **out)


def nnls_fixture():
    """Per-row proximal NNLS with Zt = Z + lam I, wt = w + lam a0 (NesterovMNLS.hpp:205-206)."""
    rng = np.random.default_rng(5)
    I, R = 7, 4
    B = rng.random((30, R))
    Z = B.T @ B
    A0 = rng.random((I, R))
    W = rng.standard_normal((I, R)) * 5.0 + 3.0
    w = np.linalg.eigvalsh(Z)
    L, mu = w.max(), w.min()
    q = mu / L
    lam = 10 * mu if 1 / q > 1e6 else (mu if 1 / q > 1e3 else mu / 10)
    Zt = Z + lam * np.eye(R)
    Wt = W + lam * A0
    C = scipy.linalg.cholesky(Zt)                      # Zt = C^T C
    sol = np.zeros((I, R))
    for i in range(I):
        d = scipy.linalg.solve_triangular(C, Wt[i], trans="T")   # C^T d = wt
        sol[i], _ = scipy.optimize.nnls(C, d)
    np.savez(os.path.join(HERE, "nnls_prox.npz"), Z=Z, W=W, A0=A0, lam=lam, sol=sol)


def philox_kat():
    kat = [
        {"ctr": [0, 0, 0, 0], "key": [0, 0], "out": [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]},
        {"ctr": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2, "out": [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]},
        {"ctr": [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], "key": [0xA4093822, 0x299F31D0],
         "out": [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]},
    ]
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)


if __name__ == "__main__":
    mttkrp_fixture([3, 4, 5], 2, 1)
    mttkrp_fixture([2, 3, 4, 5], 3, 2)
    unfold_fixture([3, 4, 5], 2)
    unfold_fixture([2, 3, 4, 5], 3)
    eig_fixture()
    nnls_fixture()
    philox_kat()
