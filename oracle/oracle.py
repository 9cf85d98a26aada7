"""ctypes binding of the FP64 CPU oracle (libptn_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and by
the CPU-baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product path.

Conventions follow the reference (Tensor.hpp:42,63-64): tensors are numpy
arrays in Fortran order (mode 0 fastest); factors are (I_n, R) arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PTN_ORACLE_LIB selects another build of the same source (the sanitised one under test)
LIB_PATH = os.environ.get("PTN_ORACLE_LIB") or os.path.join(_HERE, "libptn_oracle.so")
MAX_ORDER = 8

UNCONSTRAINED, NONNEGATIVITY, ORTHOGONALITY, SPARSITY, CONSTANT = 0, 1, 2, 3, 4
MTTKRP_STANDARD, MTTKRP_BLOCKED, MTTKRP_DIMTREE = 0, 1, 2
COST_EXPLICIT, COST_GRAM = 0, 1

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_pd = ctypes.POINTER(ctypes.c_double)
_pi64 = ctypes.POINTER(ctypes.c_int64)


class Options(ctypes.Structure):
    _fields_ = [
        ("max_iter", _i), ("threshold_error", _d), ("delta1", _d), ("delta2", _d),
        ("acceleration", _i), ("accel_coeff", _i), ("accel_fail", _i), ("normalization", _i),
        ("constraints", _i * MAX_ORDER), ("mttkrp_form", _i), ("cost_form", _i),
        ("threads", _i), ("max_inner", _i),
        ("force_accept", ctypes.POINTER(ctypes.c_int)), ("force_n", _i),
    ]


class Status(ctypes.Structure):
    _fields_ = [("frob_sq", _d), ("f_value", _d), ("rel_cost", _d), ("ao_iter", _i),
                ("accel_coeff", _i), ("accel_fail", _i)]


class SweepLog(ctypes.Structure):
    _fields_ = [
        ("f_value", _d), ("rel_cost", _d), ("f_accel", _d), ("accel_tried", _i), ("accel_accepted", _i),
        ("accel_overridden", _i), ("nnls_iters", _i * MAX_ORDER), ("zero_col_fallback", _i * MAX_ORDER),
        ("lipschitz", _d * MAX_ORDER), ("mu", _d * MAX_ORDER), ("t_start", _d), ("t_end", _d),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.ptno_uniform.restype = _d
        L.ptno_uniform.argtypes = [_u64, _u32, _u64]
        L.ptno_normal.restype = _d
        L.ptno_normal.argtypes = [_u64, _u32, _u64]
        L.ptno_exponential.restype = _d
        L.ptno_exponential.argtypes = [_u64, _u32, _u64]
        L.ptno_add_noise.restype = _d
        L.ptno_square_norm.restype = _d
        L.ptno_update_alpha.restype = _d
        L.ptno_update_alpha.argtypes = [_d, _d]
        L.ptno_cost_explicit.restype = _d
        L.ptno_cost_gram.restype = _d
        L.ptno_cost_gram.argtypes = [_i64, _i, _d, _pd, _pd, _pd, _pd]
        L.ptno_add_noise.argtypes = [_i64, _pd, _u64, _d]
        L.ptno_square_norm.argtypes = [_i64, _pd]
        L.ptno_glambda.argtypes = [_d, _pd, _pd, _pd]
        L.ptno_nesterov_mnls.argtypes = [_i64, _i, _pd, _pd, _d, _d, _i, _pd, _pd, _pd]
        L.ptno_gram.argtypes = [_i64, _i, _pd, _pd]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_pd)


def _dims(dims):
    return (ctypes.c_int64 * len(dims))(*[int(x) for x in dims])


def pack(factors):
    """Concatenate (I_n, R) factors, each column-major, factor 0 first."""
    return np.ascontiguousarray(np.concatenate([np.asarray(f, dtype=np.float64).ravel(order="F") for f in factors]))


def unpack(packed, dims, R):
    out, o = [], 0
    for I in dims:
        out.append(np.array(packed[o:o + I * R].reshape((I, R), order="F")))
        o += I * R
    return out


def _flat(X):
    return np.ascontiguousarray(np.asarray(X, dtype=np.float64).ravel(order="F"))


def philox(ctr, key):
    c = (_u32 * 4)(*ctr)
    k = (_u32 * 2)(*key)
    o = (_u32 * 4)()
    lib().ptno_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def uniform(seed, stream, idx):
    return lib().ptno_uniform(seed, stream, idx)


def normal(seed, stream, idx):
    return lib().ptno_normal(seed, stream, idx)


def gen_factors(dims, R, seed, dist=0):
    n = sum(dims) * R
    buf = np.empty(n, dtype=np.float64)
    lib().ptno_gen_factors(_i(len(dims)), _dims(dims), _i(R), _u64(seed), _i(dist), _p(buf))
    return unpack(buf, dims, R)


def cp_full(factors):
    dims = [f.shape[0] for f in factors]
    R = factors[0].shape[1]
    pk = pack(factors)
    X = np.empty(int(np.prod(dims)), dtype=np.float64)
    lib().ptno_cp_full(_i(len(dims)), _dims(dims), _i(R), _p(pk), _p(X))
    return X.reshape(dims, order="F")


def add_noise(X, seed, snr_db):
    flat = _flat(X).copy()
    s = lib().ptno_add_noise(flat.size, _p(flat), seed, snr_db)
    return flat.reshape(X.shape, order="F"), s


def square_norm(X):
    flat = _flat(X)
    return lib().ptno_square_norm(flat.size, _p(flat))


def make_tensor(dims, R, seed, snr_db=None, dist=0):
    """Config-style synthetic tensor: [[A]] (seed) plus Gaussian noise at snr_db."""
    A = gen_factors(dims, R, seed, dist)
    X = cp_full(A)
    if snr_db is not None:
        X, _ = add_noise(X, seed, snr_db)
    return X, A


def unfold(X, mode):
    dims = X.shape
    flat = _flat(X)
    out = np.empty(flat.size, dtype=np.float64)
    lib().ptno_unfold(_i(len(dims)), _dims(dims), _p(flat), _i(mode), _p(out))
    return out.reshape((dims[mode], flat.size // dims[mode]), order="F")


def partial_khatri_rao(factors, mode):
    dims = [f.shape[0] for f in factors]
    R = factors[0].shape[1]
    J = int(np.prod(dims)) // dims[mode]
    out = np.empty(J * R, dtype=np.float64)
    lib().ptno_partial_khatri_rao(_i(len(dims)), _dims(dims), _i(R), _p(pack(factors)), _i(mode), _p(out))
    return out.reshape((J, R), order="F")


def mttkrp(X, factors, mode, form="naive", threads=0):
    dims = X.shape
    R = factors[0].shape[1]
    flat = _flat(X)
    M = np.empty(dims[mode] * R, dtype=np.float64)
    if form == "naive":
        lib().ptno_mttkrp_naive(_i(len(dims)), _dims(dims), _i(R), _p(flat), _p(pack(factors)), _i(mode), _p(M))
    else:
        lib().ptno_mttkrp_blocked(_i(len(dims)), _dims(dims), _i(R), _p(flat), _p(pack(factors)), _i(mode),
                                  _i(threads), _p(M))
    return M.reshape((dims[mode], R), order="F")


def mttkrp_dimtree_all(X, factors):
    dims = X.shape
    R = factors[0].shape[1]
    out = np.empty(sum(dims) * R, dtype=np.float64)
    rc = lib().ptno_mttkrp_dimtree_all(_i(len(dims)), _dims(dims), _i(R), _p(_flat(X)), _p(pack(factors)), _p(out))
    if rc != 0:
        raise ValueError("dimension tree supports N = 3, 4")
    return unpack(out, dims, R)


def gram(A):
    A = np.asfortranarray(A, dtype=np.float64)
    R = A.shape[1]
    G = np.empty((R, R), dtype=np.float64, order="F")
    lib().ptno_gram(A.shape[0], R, _p(A), _p(G))
    return G


def partial_cwise(grams, mode):
    R = grams[0].shape[0]
    pk = np.ascontiguousarray(np.concatenate([np.asarray(g, dtype=np.float64).ravel(order="F") for g in grams]))
    V = np.empty((R, R), dtype=np.float64, order="F")
    lib().ptno_partial_cwise(_i(len(grams)), _i(R), _p(pk), _i(mode), _p(V))
    return V


def sym_extremes(V):
    V = np.asfortranarray(V, dtype=np.float64)
    L, mu = _d(), _d()
    sweeps = lib().ptno_sym_extremes(_i(V.shape[0]), _p(V), ctypes.byref(L), ctypes.byref(mu))
    return L.value, mu.value, sweeps


def sym_eig(V):
    V = np.asfortranarray(V, dtype=np.float64)
    R = V.shape[0]
    ev = np.empty(R)
    Q = np.empty((R, R), order="F")
    lib().ptno_sym_eig(_i(R), _p(V), _p(ev), _p(Q))
    return ev, Q


def glambda(mu, L):
    Lc, lam, q = _d(L), _d(), _d()
    lib().ptno_glambda(mu, ctypes.byref(Lc), ctypes.byref(lam), ctypes.byref(q))
    return Lc.value, lam.value, q.value


def update_alpha(alpha, q):
    return lib().ptno_update_alpha(alpha, q)


def nesterov(Z, W, A0, d1=1e-2, d2=1e-2, max_inner=50):
    Z = np.asfortranarray(Z, dtype=np.float64)
    W = np.asfortranarray(W, dtype=np.float64)
    A = np.array(A0, dtype=np.float64, order="F")
    L, mu = _d(), _d()
    it = lib().ptno_nesterov_mnls(A.shape[0], A.shape[1], _p(Z), _p(W), d1, d2, max_inner, _p(A),
                                  ctypes.byref(L), ctypes.byref(mu))
    return A, it, L.value, mu.value


def inverse(V):
    V = np.asfortranarray(V, dtype=np.float64)
    out = np.empty_like(V, order="F")
    rc = lib().ptno_inverse(_i(V.shape[0]), _p(V), _p(out))
    if rc != 0:
        raise np.linalg.LinAlgError("singular")
    return out


def update_factor(constraint, V, M, A0, d1=1e-2, d2=1e-2, max_inner=50):
    """update_factor (Cpd.hpp:189-228) on one factor: returns (A, G, nesterov steps, fallback)."""
    V = np.asfortranarray(V, dtype=np.float64)
    M = np.asfortranarray(M, dtype=np.float64)
    A = np.array(A0, dtype=np.float64, order="F")
    G = np.empty_like(V, order="F")
    fb = _i()
    it = lib().ptno_update_factor(_i(constraint), _i64(A.shape[0]), _i(A.shape[1]), _p(V), _p(M), _d(d1), _d(d2),
                                  _i(max_inner), _p(A), _p(G), ctypes.byref(fb))
    return A, G, it, fb.value


def normalize(grams, factors, weight_factor=0):
    dims = [f.shape[0] for f in factors]
    R = factors[0].shape[1]
    gp = np.ascontiguousarray(np.concatenate([np.asarray(g, dtype=np.float64).ravel(order="F") for g in grams]))
    fp = pack(factors)
    lib().ptno_normalize(_i(len(dims)), _dims(dims), _i(R), _i(weight_factor), _p(gp), _p(fp))
    gs = [np.array(gp[k * R * R:(k + 1) * R * R].reshape((R, R), order="F")) for k in range(len(dims))]
    return gs, unpack(fp, dims, R)


def cost_explicit(X, factors):
    dims = X.shape
    R = factors[0].shape[1]
    return lib().ptno_cost_explicit(_i(len(dims)), _dims(dims), _i(R), _p(_flat(X)), _p(pack(factors)))


def cost_gram(frob_sq, M_last, A_last, V_last, G_last):
    M = np.asfortranarray(M_last, dtype=np.float64)
    A = np.asfortranarray(A_last, dtype=np.float64)
    V = np.asfortranarray(V_last, dtype=np.float64)
    G = np.asfortranarray(G_last, dtype=np.float64)
    return lib().ptno_cost_gram(M.shape[0], M.shape[1], frob_sq, _p(M), _p(A), _p(V), _p(G))


def default_options():
    o = Options()
    lib().ptno_default_options(ctypes.byref(o))
    return o


def make_options(**kw):
    o = default_options()
    for k, v in kw.items():
        if k == "constraints":
            for i, c in enumerate(v):
                o.constraints[i] = int(c)
        else:
            setattr(o, k, v)
    return o


def _log_to_dict(rec, N):
    return {
        "f_value": rec.f_value, "rel_cost": rec.rel_cost, "f_accel": rec.f_accel,
        "accel_tried": rec.accel_tried, "accel_accepted": rec.accel_accepted,
        "accel_overridden": rec.accel_overridden, "nnls_iters": list(rec.nnls_iters[:N]), "zero_col_fallback": list(rec.zero_col_fallback[:N]),
        "lipschitz": list(rec.lipschitz[:N]), "mu": list(rec.mu[:N]),
        "t_start": rec.t_start, "t_end": rec.t_end,
    }


def cpd(X, R, init, **kw):
    """ALS as cpd()/cpd_omp/cpdDimTree (Cpd.hpp:288-356, CpdOpenMP.hpp:266-329,
    CpdDimTree.hpp:317-401), chosen by mttkrp_form/cost_form.  Returns
    (factors, status dict, per-sweep log)."""
    dims = X.shape
    N = len(dims)
    force = kw.pop("force_accept", None)
    opt = make_options(**kw)
    if force is not None:
        # forced line-search outcomes per sweep (1 accept, 0 reject, -1 the oracle's own test)
        fbuf = (ctypes.c_int * max(len(force), 1))(*[int(v) for v in force])
        opt.force_accept = ctypes.cast(fbuf, ctypes.POINTER(ctypes.c_int))
        opt.force_n = len(force)
    out = np.empty(sum(dims) * R, dtype=np.float64)
    st = Status()
    cap = max(int(opt.max_iter), 1)
    logs = (SweepLog * cap)()
    rc = lib().ptno_cpd(_i(N), _dims(dims), _i(R), _p(_flat(X)), _p(pack(init)), ctypes.byref(opt), _p(out),
                        ctypes.byref(st), logs, _i(cap))
    if rc != 0:
        raise ValueError("ptno_cpd failed: %d" % rc)
    status = {"frob_sq": st.frob_sq, "f_value": st.f_value, "rel_cost": st.rel_cost, "ao_iter": st.ao_iter,
              "accel_coeff": st.accel_coeff, "accel_fail": st.accel_fail}
    return unpack(out, dims, R), status, [_log_to_dict(logs[i], N) for i in range(min(st.ao_iter, cap))]
