"""Test infrastructure: FP64 NumPy restatement of the NNLS warm-start network.

Only tests/, smoke() and bench.py's CPU leg use this file, as the checker of the HIP
path.  The network has no counterpart in the reference (NesterovMNLS.hpp:181-234 always
starts from the previous factor); its definition here is the mathematics of the library's
kernels: the row features, a 2208 -> 256 -> 256 -> 64 ReLU MLP, the masked mean squared
error and its gradient by back-propagation.  Pinned by central finite differences in FP64
(tests/test_mlp_oracle.py).  `bf16=True` rounds what the device stores in BF16 (inputs,
weights, hidden activations, output gradient) so that device and oracle see the same
operands.
"""
import numpy as np

R_MAX = 64
VECH = R_MAX * (R_MAX + 1) // 2
IN = VECH + 2 * R_MAX
H = 256
N_PARAMS = H * IN + H + H * H + H + R_MAX * H + R_MAX


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def vech_index(a, b):
    """Position of entry (a, b), a <= b, in the row-by-row upper triangle of a 64 x 64 matrix."""
    return a * R_MAX - (a * (a - 1)) // 2 + (b - a)


def features(V, M, A, Anew=None):
    """Rows of the network's input for one NNLS update: V (R x R), M and A (I x R, the
    MTTKRP and the previous factor).  sV = max diag(V); s_m = max_c |m_c| per row (1 if 0).
    Returns X (I x 2208), the targets Anew * sV / s_m (I x 64, zero beyond R) when Anew is
    given, and the per-row scale s_m / sV that maps outputs back."""
    V, M, A = np.asarray(V, float), np.asarray(M, float), np.asarray(A, float)
    R = V.shape[0]
    I = M.shape[0]
    sv = np.max(np.diag(V))
    sv = sv if sv > 0 else 1.0
    sm = np.max(np.abs(M), axis=1)
    sm = np.where(sm > 0, sm, 1.0)
    X = np.zeros((I, IN))
    iu = np.triu_indices(R)
    pos = np.array([vech_index(a, b) for a, b in zip(*iu)], dtype=np.int64)
    X[:, pos] = (V[iu] / sv)[None, :]
    X[:, VECH:VECH + R] = M / sm[:, None]
    X[:, VECH + R_MAX:VECH + R_MAX + R] = A * sv / sm[:, None]
    T = None
    if Anew is not None:
        T = np.zeros((I, R_MAX))
        T[:, :R] = np.asarray(Anew, float) * sv / sm[:, None]
    return X, T, sm / sv


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    out = []
    for shape in ((H, IN), (H,), (H, H), (H,), (R_MAX, H), (R_MAX,)):
        n = int(np.prod(shape))
        out.append(w[o:o + n].reshape(shape))
        o += n
    return out


def forward(w, X, bf16=False):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    X = np.asarray(X, float)
    if bf16:
        X, W1, W2, W3 = bf16_round(X), bf16_round(W1), bf16_round(W2), bf16_round(W3)
    H1 = np.maximum(X @ W1.T + b1, 0.0)
    if bf16:
        H1 = bf16_round(H1)
    H2 = np.maximum(H1 @ W2.T + b2, 0.0)
    if bf16:
        H2 = bf16_round(H2)
    Y = H2 @ W3.T + b3
    return Y, (X, W1, W2, W3, H1, H2)


def loss_grad(w, X, T, rank, bf16=False):
    """Masked MSE (entries c < rank[i]) averaged over the valid entries, and its gradient
    in the flat parameter layout."""
    Y, (Xr, W1, W2, W3, H1, H2) = forward(w, X, bf16)
    rank = np.asarray(rank)
    mask = (np.arange(R_MAX)[None, :] < rank[:, None]).astype(float)
    cnt = mask.sum()
    D = (Y - np.asarray(T, float)) * mask
    loss = float((D * D).sum() / cnt)
    dY = 2.0 * D / cnt
    dYg = bf16_round(dY) if bf16 else dY
    gW3 = dYg.T @ H2
    gb3 = dY.sum(axis=0)
    dH2 = (dYg @ W3) * (H2 > 0)
    if bf16:
        dH2 = bf16_round(dH2)
    gW2 = dH2.T @ H1
    gb2 = dH2.sum(axis=0)
    dH1 = (dH2 @ W2) * (H1 > 0)
    if bf16:
        dH1 = bf16_round(dH1)
    gW1 = dH1.T @ Xr
    gb1 = dH1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def kkt_residual(V, lam, L, M, Aprev, A):
    """Relative KKT residual of start rows A for the subproblem NesterovMNLS solves
    (Zt = V + lam I, Wt = M + lam Aprev, Lipschitz constant L): the projected-gradient
    fixed-point residual min(a, g / L), g = a Zt - Wt, relative to max|Wt| / L, i.e.
    max_c |min(L a_c, g_c)| / max_c |Wt_c| per row."""
    R = V.shape[0]
    Zt = V + lam * np.eye(R)
    Wt = M + lam * Aprev
    G = A @ Zt - Wt
    den = np.max(np.abs(Wt), axis=1)
    den = np.where(den > 0, den, 1.0)
    return np.max(np.abs(np.minimum(L * A, G)), axis=1) / den
