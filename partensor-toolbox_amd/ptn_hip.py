"""Python binding of libptn_hip.so (the C ABI declared in include/ptn_hip.h).

Mirrors the reference's ALS interface for this path: options with the
Options<Tensor> fields (PARTENSOR_basic.hpp:283-336), a status with the
Status<Tensor> fields (PARTENSOR_basic.hpp:556-595), and cpd() / cpd_dimtree()
taking a tensor, a rank, options and initial factors (Cpd.hpp:550-577,
CpdDimTree.hpp:534-565).  Tensors are numpy arrays in Fortran order (mode 0
fastest), factors (I_n, R) arrays.  There is no CPU fallback: if the HIP
library is missing or no GPU is present the calls raise.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libptn_hip.so")
MAX_ORDER = 8
MAX_RANK = 64

UNCONSTRAINED, NONNEGATIVITY, ORTHOGONALITY, SPARSITY, CONSTANT = 0, 1, 2, 3, 4
TENSOR_F64, TENSOR_F32, TENSOR_BOTH = 1, 2, 3
MTTKRP_STANDARD, MTTKRP_DIMTREE = 0, 2
ARITH_F64, ARITH_F32, ARITH_BF16X3, ARITH_F32_F64 = 0, 1, 2, 3

EXPORTED = [
    "ptn_create", "ptn_destroy", "ptn_last_error", "ptn_default_options", "ptn_abi_version",
    "ptn_set_tensor", "ptn_generate_tensor", "ptn_generate_factors", "ptn_tensor_norm_sq",
    "ptn_cpd", "ptn_mttkrp", "ptn_mttkrp_dimtree", "ptn_nesterov_mnls", "ptn_nesterov_mnls_noalloc",
    "ptn_sym_extremes",
    "ptn_normalize", "ptn_cost_explicit", "ptn_time_contraction",
    "ptn_batch_set", "ptn_batch_generate", "ptn_batch_init_factors", "ptn_batch_get_tensor", "ptn_cpd_batch",
    "ptn_mlp_init", "ptn_mlp_set_params", "ptn_mlp_get_params", "ptn_mlp_forward", "ptn_mlp_loss_grad",
    "ptn_mlp_data_clear", "ptn_mlp_data_add", "ptn_mlp_data_size", "ptn_mlp_data_get", "ptn_mlp_train",
]
MLP_IN, MLP_OUT, MLP_PARAMS = 2208, 64, 647744

_d, _i, _i64, _u64 = ctypes.c_double, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p


class Options(ctypes.Structure):
    _fields_ = [
        ("max_iter", _i), ("threshold_error", _d), ("nesterov_delta_1", _d), ("nesterov_delta_2", _d),
        ("accel_coeff", _i), ("accel_fail", _i), ("acceleration", _i), ("normalization", _i),
        ("constraints", _i * MAX_ORDER), ("mttkrp_form", _i), ("arith", _i), ("max_inner", _i),
        ("fuse_accel", _i), ("warm_start", _i), ("warm_kkt_tol", _d), ("warm_refine", _i), ("mlp_harvest", _i), ("time_passes", _i),
    ]


class Result(ctypes.Structure):
    _fields_ = [("frob_sq", _d), ("f_value", _d), ("rel_cost", _d), ("ao_iter", _i), ("accel_coeff", _i),
                ("accel_fail", _i), ("sweep_seconds", _d)]


class SweepLog(ctypes.Structure):
    _fields_ = [
        ("f_value", _d), ("rel_cost", _d), ("f_accel", _d), ("accel_tried", _i), ("accel_accepted", _i),
        ("nnls_iters", _i * MAX_ORDER), ("zero_col_fallback", _i * MAX_ORDER),
        ("lipschitz", _d * MAX_ORDER), ("mu", _d * MAX_ORDER), ("t_end", _d), ("warm_rows", _i * MAX_ORDER),
        ("t_pass_ms", _d * 2), ("warm_steps", _i * MAX_ORDER),
    ]


class BatchLog(ctypes.Structure):
    _fields_ = [("f_value", _d), ("f_accel", _d), ("nnls_iters", _i * 3), ("accel_tried", _i),
                ("accel_accepted", _i), ("stopped", _i)]


class BatchReport(ctypes.Structure):
    _fields_ = [("block", _i), ("grid", _i), ("tuned", _i), ("sweeps", _i), ("seconds", _d),
                ("tune_seconds", _d), ("split_problems", _i), ("max_pieces", _i)]


class PtnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ptn status %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    """Load libptn_hip.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libptn_hip.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.ptn_create.argtypes = [ctypes.POINTER(_vp), _i]
        L.ptn_destroy.argtypes = [_vp]
        L.ptn_destroy.restype = None
        L.ptn_last_error.argtypes = [_vp]
        L.ptn_last_error.restype = ctypes.c_char_p
        L.ptn_default_options.argtypes = [ctypes.POINTER(Options)]
        L.ptn_default_options.restype = None
        L.ptn_set_tensor.argtypes = [_vp, _i, _pi64, _pd, _i]
        L.ptn_generate_tensor.argtypes = [_vp, _i, _pi64, _i, _u64, _i, _d, _i]
        L.ptn_generate_factors.argtypes = [_vp, _i, _pi64, _i, _u64, _i, _pd]
        L.ptn_tensor_norm_sq.argtypes = [_vp, _pd]
        L.ptn_cpd.argtypes = [_vp, _i, _pd, ctypes.POINTER(Options), _pd, ctypes.POINTER(Result),
                              ctypes.POINTER(SweepLog), _i]
        L.ptn_mttkrp.argtypes = [_vp, _i, _pd, _i, _i, _pd]
        L.ptn_mttkrp_dimtree.argtypes = [_vp, _i, _pd, _i, _pd]
        L.ptn_nesterov_mnls.argtypes = [_vp, _i64, _i, _pd, _pd, _d, _d, _i, _pd, ctypes.POINTER(_i), _pd, _pd]
        L.ptn_nesterov_mnls_noalloc.argtypes = [_vp, _i64, _i, _vp, _vp, _d, _d, _i, _vp, ctypes.POINTER(_i), _pd,
                                                _pd]
        L.ptn_sym_extremes.argtypes = [_vp, _i, _pd, _pd, _pd]
        L.ptn_normalize.argtypes = [_vp, _i, _pi64, _i, _i, _pd, _pd]
        L.ptn_cost_explicit.argtypes = [_vp, _i, _pd, _pd]
        L.ptn_time_contraction.argtypes = [_vp, _i, _i, _pd, _i, _i, _pd, _pd]
        L.ptn_batch_set.argtypes = [_vp, _i, _pi, _pi, _pd]
        L.ptn_batch_generate.argtypes = [_vp, _i, _pi, _pi, _u64, _d]
        L.ptn_batch_init_factors.argtypes = [_vp, _u64, _pd]
        L.ptn_batch_get_tensor.argtypes = [_vp, _i, _pd]
        L.ptn_cpd_batch.argtypes = [_vp, _pd, ctypes.POINTER(Options), _pd, ctypes.POINTER(Result),
                                    ctypes.POINTER(BatchLog), _i, _i, ctypes.POINTER(BatchReport)]
        _pf = ctypes.POINTER(ctypes.c_float)
        _pu8 = ctypes.POINTER(ctypes.c_ubyte)
        L.ptn_mlp_init.argtypes = [_vp, _u64]
        L.ptn_mlp_set_params.argtypes = [_vp, _pf]
        L.ptn_mlp_get_params.argtypes = [_vp, _pf]
        L.ptn_mlp_forward.argtypes = [_vp, _i, _pf, _pf]
        L.ptn_mlp_loss_grad.argtypes = [_vp, _i, _pf, _pf, _pu8, _pd, _pf]
        L.ptn_mlp_data_clear.argtypes = [_vp]
        L.ptn_mlp_data_add.argtypes = [_vp, _i, _pf, _pf, _pu8]
        L.ptn_mlp_data_size.argtypes = [_vp, ctypes.POINTER(_i64)]
        L.ptn_mlp_train.argtypes = [_vp, _i, _i, _d, _u64, _pd]
        L.ptn_mlp_data_get.argtypes = [_vp, _i64, _i, _pf, _pf, _pu8]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_pd)


def _pf32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _pu8(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))


def _dims(dims):
    return (ctypes.c_int64 * len(dims))(*[int(x) for x in dims])


def pack(factors):
    return np.ascontiguousarray(np.concatenate([np.asarray(f, dtype=np.float64).ravel(order="F") for f in factors]))


def unpack(packed, dims, R):
    out, o = [], 0
    for I in dims:
        out.append(np.array(packed[o:o + I * R].reshape((I, R), order="F")))
        o += I * R
    return out


def default_options(**kw):
    o = Options()
    lib().ptn_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if k == "constraints":
            for i, c in enumerate(v):
                o.constraints[i] = int(c)
        else:
            setattr(o, k, v)
    return o


class Handle:
    """Owns one device context (streams, tensor, work buffers) through the C ABI."""

    def __init__(self, device=0):
        self._h = _vp()
        rc = lib().ptn_create(ctypes.byref(self._h), device)
        if rc != 0:
            raise PtnError(rc, "ptn_create failed (device %d)" % device)
        self.dims = None

    def close(self):
        if self._h:
            lib().ptn_destroy(self._h)
            self._h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise PtnError(rc, lib().ptn_last_error(self._h).decode())

    # -- tensor --
    def set_tensor(self, X, storage=TENSOR_F64):
        X = np.asarray(X, dtype=np.float64)
        flat = np.ascontiguousarray(X.ravel(order="F"))
        self._check(lib().ptn_set_tensor(self._h, X.ndim, _dims(X.shape), _p(flat), storage))
        self.dims = list(X.shape)

    def generate_tensor(self, dims, R_true, seed, snr_db=None, dist=0, storage=TENSOR_F64):
        snr = float("nan") if snr_db is None else float(snr_db)
        self._check(lib().ptn_generate_tensor(self._h, len(dims), _dims(dims), R_true, seed, dist, snr, storage))
        self.dims = list(dims)

    def generate_factors(self, dims, R, seed, dist=0):
        out = np.empty(sum(dims) * R, dtype=np.float64)
        self._check(lib().ptn_generate_factors(self._h, len(dims), _dims(dims), R, seed, dist, _p(out)))
        return unpack(out, dims, R)

    def norm_sq(self):
        v = _d()
        self._check(lib().ptn_tensor_norm_sq(self._h, ctypes.byref(v)))
        return v.value

    # -- ALS --
    def cpd(self, R, init, options=None, **kw):
        """cpd(tnsX, R, options, factorsInit) on the handle's tensor.  Returns
        (factors, status dict, per-sweep log list)."""
        o = options if options is not None else default_options(**kw)
        dims = self.dims
        out = np.empty(sum(dims) * R, dtype=np.float64)
        res = Result()
        cap = max(int(o.max_iter), 1)
        logs = (SweepLog * cap)()
        self._check(lib().ptn_cpd(self._h, R, _p(pack(init)), ctypes.byref(o), _p(out), ctypes.byref(res), logs,
                                  cap))
        N = len(dims)
        status = {"frob_sq": res.frob_sq, "f_value": res.f_value, "rel_cost": res.rel_cost, "ao_iter": res.ao_iter,
                  "accel_coeff": res.accel_coeff, "accel_fail": res.accel_fail,
                  "sweep_seconds": res.sweep_seconds}
        log = []
        for i in range(min(res.ao_iter, cap)):
            r = logs[i]
            log.append({"f_value": r.f_value, "rel_cost": r.rel_cost, "f_accel": r.f_accel,
                        "accel_tried": r.accel_tried, "accel_accepted": r.accel_accepted,
                        "nnls_iters": list(r.nnls_iters[:N]), "zero_col_fallback": list(r.zero_col_fallback[:N]),
                        "lipschitz": list(r.lipschitz[:N]), "mu": list(r.mu[:N]), "t_end": r.t_end,
                        "warm_rows": list(r.warm_rows[:N]), "warm_steps": list(r.warm_steps[:N]),
                        "t_pass_ms": list(r.t_pass_ms)})
        return unpack(out, dims, R), status, log

    def cpd_dimtree(self, R, init, **kw):
        kw.setdefault("mttkrp_form", MTTKRP_DIMTREE)
        return self.cpd(R, init, **kw)

    # -- the NNLS warm-start network --
    def mlp_init(self, seed):
        self._check(lib().ptn_mlp_init(self._h, seed))

    def mlp_params(self):
        w = np.empty(MLP_PARAMS, dtype=np.float32)
        self._check(lib().ptn_mlp_get_params(self._h, _pf32(w)))
        return w

    def mlp_set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == MLP_PARAMS
        self._check(lib().ptn_mlp_set_params(self._h, _pf32(w)))

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], MLP_OUT), dtype=np.float32)
        self._check(lib().ptn_mlp_forward(self._h, X.shape[0], _pf32(X), _pf32(Y)))
        return Y

    def mlp_loss_grad(self, X, T, rank):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        rk = np.ascontiguousarray(rank, dtype=np.uint8)
        loss = _d()
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        self._check(lib().ptn_mlp_loss_grad(self._h, X.shape[0], _pf32(X), _pf32(T), _pu8(rk), ctypes.byref(loss),
                                            _pf32(g)))
        return loss.value, g

    def mlp_data_clear(self):
        self._check(lib().ptn_mlp_data_clear(self._h))

    def mlp_data_add(self, X, T, rank):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        rk = np.ascontiguousarray(rank, dtype=np.uint8)
        self._check(lib().ptn_mlp_data_add(self._h, X.shape[0], _pf32(X), _pf32(T), _pu8(rk)))

    def mlp_data_size(self):
        n = _i64()
        self._check(lib().ptn_mlp_data_size(self._h, ctypes.byref(n)))
        return n.value

    def mlp_data_get(self, first, n):
        X = np.empty((n, MLP_IN), dtype=np.float32)
        T = np.empty((n, MLP_OUT), dtype=np.float32)
        rk = np.empty(n, dtype=np.uint8)
        self._check(lib().ptn_mlp_data_get(self._h, first, n, _pf32(X), _pf32(T), _pu8(rk)))
        return X, T, rk

    def mlp_train(self, epochs, batch=1024, lr=1e-3, seed=0):
        out = np.empty(epochs, dtype=np.float64)
        self._check(lib().ptn_mlp_train(self._h, epochs, batch, lr, seed, _p(out)))
        return out

    # -- batched independent order-3 problems --
    def _set_batch_shape(self, dims, ranks):
        d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1, 3))
        r = np.ascontiguousarray(np.asarray(ranks, dtype=np.int32).ravel())
        if d.shape[0] != r.shape[0]:
            raise ValueError("dims and ranks describe different numbers of problems")
        self.batch_dims, self.batch_ranks = [tuple(int(v) for v in x) for x in d], [int(v) for v in r]
        return d, r

    def batch_set(self, tensors, ranks):
        """Upload a batch: a list of order-3 arrays (stored FP32 on the device)."""
        dims = [t.shape for t in tensors]
        d, r = self._set_batch_shape(dims, ranks)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.float64).ravel(order="F")
                                                    for t in tensors]))
        self._check(lib().ptn_batch_set(self._h, d.shape[0], d.ctypes.data_as(_pi), r.ctypes.data_as(_pi),
                                        _p(flat)))

    def batch_generate(self, dims, ranks, seed, snr_db=None):
        d, r = self._set_batch_shape(dims, ranks)
        snr = float("nan") if snr_db is None else float(snr_db)
        self._check(lib().ptn_batch_generate(self._h, d.shape[0], d.ctypes.data_as(_pi), r.ctypes.data_as(_pi),
                                             seed, snr))

    def batch_init_factors(self, seed):
        """Initial factors of every problem: a list (per problem) of [A0, A1, A2]."""
        n = sum(sum(dd) * R for dd, R in zip(self.batch_dims, self.batch_ranks))
        out = np.empty(n, dtype=np.float64)
        self._check(lib().ptn_batch_init_factors(self._h, seed, _p(out)))
        return self._unpack_batch(out)

    def batch_tensor(self, p):
        dd = self.batch_dims[p]
        out = np.empty(dd[0] * dd[1] * dd[2], dtype=np.float64)
        self._check(lib().ptn_batch_get_tensor(self._h, p, _p(out)))
        return out.reshape(dd, order="F")

    def _unpack_batch(self, flat):
        res, o = [], 0
        for dd, R in zip(self.batch_dims, self.batch_ranks):
            n = sum(dd) * R
            res.append(unpack(flat[o:o + n], dd, R))
            o += n
        return res

    def cpd_batch(self, init, options=None, block=0, grid=0, with_log=True, **kw):
        """ALS on every problem of the batch.  Returns (factors per problem, status dict per
        problem, per-problem sweep logs or None, report dict)."""
        o = options if options is not None else default_options(**kw)
        P = len(self.batch_ranks)
        flat = np.ascontiguousarray(np.concatenate([pack(f) for f in init]))
        out = np.empty_like(flat)
        res = (Result * P)()
        logs = (BatchLog * (P * int(o.max_iter)))() if with_log else None
        rep = BatchReport()
        self._check(lib().ptn_cpd_batch(self._h, _p(flat), ctypes.byref(o), _p(out), res, logs, block, grid,
                                        ctypes.byref(rep)))
        status = [{"frob_sq": r.frob_sq, "f_value": r.f_value, "rel_cost": r.rel_cost, "ao_iter": r.ao_iter,
                   "accel_coeff": r.accel_coeff, "accel_fail": r.accel_fail} for r in res]
        log = None
        if with_log:
            m = int(o.max_iter)
            log = []
            for p in range(P):
                lp = []
                for i in range(min(res[p].ao_iter, m)):
                    r = logs[p * m + i]
                    lp.append({"f_value": r.f_value, "f_accel": r.f_accel, "nnls_iters": list(r.nnls_iters),
                               "accel_tried": r.accel_tried, "accel_accepted": r.accel_accepted,
                               "stopped": r.stopped})
                log.append(lp)
        report = {"block": rep.block, "grid": rep.grid, "tuned": rep.tuned, "sweeps": rep.sweeps,
                  "seconds": rep.seconds, "tune_seconds": rep.tune_seconds, "split_problems": rep.split_problems,
                  "max_pieces": rep.max_pieces}
        return self._unpack_batch(out), status, log, report

    # -- single operations --
    def mttkrp(self, factors, mode, arith=ARITH_F64):
        R = factors[0].shape[1]
        rows = self.dims[mode] if 0 <= mode < len(self.dims) else 1   # the C ABI rejects a bad mode
        M = np.empty(rows * R, dtype=np.float64)
        self._check(lib().ptn_mttkrp(self._h, R, _p(pack(factors)), mode, arith, _p(M)))
        return M.reshape((rows, R), order="F")

    def mttkrp_dimtree(self, factors, arith=ARITH_F64):
        R = factors[0].shape[1]
        out = np.empty(sum(self.dims) * R, dtype=np.float64)
        self._check(lib().ptn_mttkrp_dimtree(self._h, R, _p(pack(factors)), arith, _p(out)))
        return unpack(out, self.dims, R)

    def nesterov(self, Z, W, A0, d1=1e-2, d2=1e-2, max_inner=50):
        Z = np.asfortranarray(Z, dtype=np.float64)
        W = np.asfortranarray(W, dtype=np.float64)
        A = np.array(A0, dtype=np.float64, order="F")
        it, L, mu = _i(), _d(), _d()
        self._check(lib().ptn_nesterov_mnls(self._h, A.shape[0], A.shape[1], _p(Z), _p(W), d1, d2, max_inner, _p(A),
                                            ctypes.byref(it), ctypes.byref(L), ctypes.byref(mu)))
        return A, it.value, L.value, mu.value

    def nesterov_noalloc(self, I, R, Z_dev, W_dev, A_dev, d1=1e-2, d2=1e-2, max_inner=50):
        """ptn_nesterov_mnls_noalloc on device buffers (addresses as ints): Z (R x R), W and A
        (I x R, A in/out), column-major.  Returns (iters, L, mu)."""
        it, L, mu = _i(), _d(), _d()
        self._check(lib().ptn_nesterov_mnls_noalloc(self._h, I, R, Z_dev, W_dev, d1, d2, max_inner, A_dev,
                                                    ctypes.byref(it), ctypes.byref(L), ctypes.byref(mu)))
        return it.value, L.value, mu.value

    def sym_extremes(self, V):
        V = np.asfortranarray(V, dtype=np.float64)
        L, mu = _d(), _d()
        self._check(lib().ptn_sym_extremes(self._h, V.shape[0], _p(V), ctypes.byref(L), ctypes.byref(mu)))
        return L.value, mu.value

    def normalize(self, grams, factors, weight_factor=0):
        dims = [f.shape[0] for f in factors]
        R = factors[0].shape[1]
        gp = np.ascontiguousarray(np.concatenate([np.asarray(g, dtype=np.float64).ravel(order="F") for g in grams]))
        fp = pack(factors)
        self._check(lib().ptn_normalize(self._h, len(dims), _dims(dims), R, weight_factor, _p(gp), _p(fp)))
        gs = [np.array(gp[k * R * R:(k + 1) * R * R].reshape((R, R), order="F")) for k in range(len(dims))]
        return gs, unpack(fp, dims, R)

    def cost_explicit(self, factors):
        R = factors[0].shape[1]
        f = _d()
        self._check(lib().ptn_cost_explicit(self._h, R, _p(pack(factors)), ctypes.byref(f)))
        return f.value

    def time_contraction(self, kernel, factors, arith=ARITH_F64, iters=10):
        R = factors[0].shape[1]
        ms, b = _d(), _d()
        self._check(lib().ptn_time_contraction(self._h, kernel, R, _p(pack(factors)), arith, iters,
                                               ctypes.byref(ms), ctypes.byref(b)))
        return ms.value, b.value


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


__all__ = ["Handle", "default_options", "pack", "unpack", "lib", "PtnError", "EXPORTED", "rel_err", "math"]
