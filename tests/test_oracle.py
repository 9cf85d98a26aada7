"""Pins of the FP64 CPU oracle (CPU only).

Each check is chosen so that a plausible mistake (dropped term, wrong index,
transposed operand, wrong branch) fails it: golden vectors from definitions,
library routines (LAPACK eigvalsh, scipy nnls), closed forms and invariants.
"""
import json
import os

import numpy as np
import pytest

import oracle as o


# --------------------------------------------------------------- generator --

def test_philox_known_answers(golden_dir):
    with open(os.path.join(golden_dir, "philox_kat.json")) as f:
        kat = json.load(f)
    for v in kat:
        assert o.philox(v["ctr"], v["key"]) == v["out"]


def test_generator_draw_layout_and_moments():
    dims, R = [7, 5, 3], 4
    A = o.gen_factors(dims, R, seed=9)
    off = 0
    for k, I in enumerate(dims):
        for i in range(I):
            for r in range(R):
                assert A[k][i, r] == o.uniform(9, 0, off + i + I * r)   # column-major, factor 0 first
        off += I * R
    u = np.array([o.uniform(3, 0, i) for i in range(20000)])
    z = np.array([o.normal(3, 1, i) for i in range(20000)])
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.005
    assert abs(z.mean()) < 0.03 and abs(z.var() - 1.0) < 0.04


def test_noise_snr():
    X, _ = o.make_tensor([20, 21, 22], 3, seed=0)
    Xn, s = o.add_noise(X, seed=0, snr_db=20.0)
    ratio = np.linalg.norm(Xn - X) / np.linalg.norm(X)
    assert abs(ratio - 10 ** (-20 / 20)) < 1e-12


def test_cp_full_matches_outer_products():
    rng = np.random.default_rng(0)
    A = [rng.random((I, 3)) for I in (4, 5, 6)]
    X = o.cp_full(A)
    ref = np.einsum("ir,jr,kr->ijk", *A)
    assert np.allclose(X, ref, rtol=1e-14, atol=0)


# ------------------------------------------------------------------ layout --

@pytest.mark.parametrize("name", ["unfold_3x4x5.npz", "unfold_2x3x4x5.npz"])
def test_unfold_and_khatri_rao_known_answer(golden_dir, name):
    g = np.load(os.path.join(golden_dir, name))
    dims = [int(d) for d in g["dims"]]
    X = g["X"].reshape(dims, order="F")
    A = [g["A%d" % k] for k in range(len(dims))]
    for mode in range(len(dims)):
        assert np.array_equal(o.unfold(X, mode), g["U%d" % mode])
        assert np.array_equal(o.partial_khatri_rao(A, mode), g["KR%d" % mode])


# ------------------------------------------------------------------ MTTKRP --

@pytest.mark.parametrize("name", ["mttkrp_3x4x5_r2.npz", "mttkrp_2x3x4x5_r3.npz"])
@pytest.mark.parametrize("form", ["naive", "blocked", "dimtree"])
def test_mttkrp_golden_exact(golden_dir, name, form):
    g = np.load(os.path.join(golden_dir, name))
    dims = [int(d) for d in g["dims"]]
    X = g["X"].reshape(dims, order="F")
    A = [g["A%d" % k] for k in range(len(dims))]
    if form == "dimtree":
        M = o.mttkrp_dimtree_all(X, A)
    for mode in range(len(dims)):
        got = M[mode] if form == "dimtree" else o.mttkrp(X, A, mode, form=form, threads=3)
        assert np.array_equal(got, g["M%d" % mode])           # integer data: exact


def test_mttkrp_equals_unfolding_times_khatri_rao():
    rng = np.random.default_rng(3)
    dims = [6, 5, 4, 3, 2]
    X = np.asfortranarray(rng.standard_normal(dims))
    A = [rng.standard_normal((I, 4)) for I in dims]
    for mode in range(len(dims)):
        ref = o.unfold(X, mode) @ o.partial_khatri_rao(A, mode)
        for form in ("naive", "blocked"):
            assert np.allclose(o.mttkrp(X, A, mode, form=form), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dims", [[9, 8, 7], [6, 5, 4, 3]])
def test_mttkrp_closed_form_low_rank(dims):
    """X = [[B]] exactly: M_n = B_n (Hadamard_{k!=n} B_k^T A_k)."""
    rng = np.random.default_rng(4)
    Rt, R = 3, 5
    B = [rng.random((I, Rt)) for I in dims]
    A = [rng.random((I, R)) for I in dims]
    X = o.cp_full(B)
    D = o.mttkrp_dimtree_all(X, A)
    for n in range(len(dims)):
        H = np.ones((Rt, R))
        for k in range(len(dims)):
            if k != n:
                H *= B[k].T @ A[k]
        ref = B[n] @ H
        assert np.allclose(o.mttkrp(X, A, n), ref, rtol=1e-12, atol=0)
        assert np.allclose(D[n], ref, rtol=1e-12, atol=0)


# ------------------------------------------------------------ R x R pieces --

def test_gram_and_partial_cwise_order():
    rng = np.random.default_rng(1)
    A = rng.random((13, 4))
    assert np.allclose(o.gram(A), A.T @ A, rtol=1e-14, atol=0)
    G = [rng.random((4, 4)) + 1 for _ in range(4)]
    for mode in range(4):
        ref = np.ones((4, 4))
        for k in range(4):
            if k != mode:
                ref = ref * G[k]
        V = o.partial_cwise(G, mode)
        assert np.allclose(V, ref, rtol=1e-15, atol=0)
        # a dropped or repeated factor changes the result
        assert not np.allclose(V, ref * G[mode])


def test_sym_extremes_vs_lapack(golden_dir):
    g = np.load(os.path.join(golden_dir, "eig_extremes.npz"))
    for i in range(int(g["count"])):
        V, w = g["V%d" % i], g["w%d" % i]
        L, mu, sweeps = o.sym_extremes(V)
        aw = np.abs(w)
        tol = 4 * V.shape[0] * 2.2e-16 * aw.max()      # backward-stable bound for both solvers
        assert abs(L - aw.max()) <= tol
        assert abs(mu - aw.min()) <= tol
        ev, Q = o.sym_eig(V)
        assert np.allclose(np.sort(ev), np.sort(w), rtol=0, atol=tol)
        assert np.allclose(Q @ np.diag(ev) @ Q.T, V, rtol=0, atol=1e-12 * aw.max())


@pytest.mark.parametrize("kappa,branch", [(2e6, 10.0), (5e4, 1.0), (50.0, 0.1)])
def test_glambda_branches(kappa, branch):
    mu, L = 2.0, 2.0 * kappa
    L2, lam, q = o.glambda(mu, L)
    assert lam == pytest.approx(branch * mu, rel=1e-15)
    assert L2 == pytest.approx(L + lam, rel=1e-15)
    assert q == pytest.approx((mu + lam) / (L + lam), rel=1e-15)


@pytest.mark.parametrize("alpha,q", [(1.0, 0.1), (0.5, 0.01), (0.2, 0.3), (1.0, 1e-6)])
def test_update_alpha_root(alpha, q):
    a = o.update_alpha(alpha, q)
    # a solves a^2 = (1 - a) alpha^2 + q a  (Nesterov's constant-step scheme, a in (0, 1])
    assert 0.0 < a <= 1.0
    assert a * a == pytest.approx((1 - a) * alpha * alpha + q * a, rel=1e-14, abs=1e-16)


def test_nesterov_converges_to_scipy_nnls(golden_dir):
    g = np.load(os.path.join(golden_dir, "nnls_prox.npz"))
    A, it, L, mu = o.nesterov(g["Z"], g["W"], g["A0"], d1=0.0, d2=0.0, max_inner=4000)
    assert np.allclose(A, g["sol"], rtol=0, atol=1e-9)
    w = np.linalg.eigvalsh(g["Z"])
    assert L == pytest.approx(w.max(), rel=1e-14) and mu == pytest.approx(w.min(), rel=1e-12)


def test_nesterov_kkt_at_exit_and_cap():
    rng = np.random.default_rng(8)
    B = rng.random((40, 6))
    Z = B.T @ B
    W = rng.standard_normal((30, 6)) * 3
    A0 = rng.random((30, 6))
    A, it, L, mu = o.nesterov(Z, W, A0, d1=1e-2, d2=1e-2, max_inner=50)
    assert (A >= 0).all()
    # delta = 0 can never be met by the projected iterate -> the 50-step cap (NesterovMNLS.hpp:217)
    A2, it2, _, _ = o.nesterov(Z, W, A0, d1=0.0, d2=0.0, max_inner=50)
    assert it2 == 50
    # huge tolerances: stop at the very first check, result = start point
    A3, it3, _, _ = o.nesterov(Z, W, A0, d1=1e30, d2=1e30, max_inner=50)
    assert it3 == 0 and np.array_equal(A3, A0)


def test_inverse():
    rng = np.random.default_rng(2)
    B = rng.random((9, 5))
    V = B.T @ B
    assert np.allclose(o.inverse(V) @ V, np.eye(5), atol=1e-10)


# --------------------------------------------------------------- Normalize --

@pytest.mark.parametrize("N", [3, 4])
def test_normalize_invariants(N):
    rng = np.random.default_rng(N)
    dims = [5, 6, 7, 4][:N]
    A = [rng.random((I, 3)) for I in dims]
    G = [a.T @ a for a in A]
    X0 = o.cp_full(A)
    G2, A2 = o.normalize(G, A, weight_factor=0)
    assert np.allclose(o.cp_full(A2), X0, rtol=1e-13, atol=0)          # [[A]] unchanged
    assert np.array_equal(A2[-1], A[-1])                               # last factor never touched
    for j in range(1, N - 1):                                          # rescaled factors: unit columns
        assert np.allclose(np.linalg.norm(A2[j], axis=0), 1.0, rtol=1e-14)
    for j in range(N):
        assert np.allclose(G2[j], A2[j].T @ A2[j], rtol=1e-12, atol=0)


def test_normalize_zero_test_reads_last_factor_only():
    rng = np.random.default_rng(0)
    A = [rng.random((4, 3)) for _ in range(3)]
    A[2][:, 1] = 0.0                         # last factor column 1 is zero
    G = [a.T @ a for a in A]
    G2, A2 = o.normalize(G, A, weight_factor=0)
    assert np.array_equal(A2[1][:, 1], A[1][:, 1])                     # column skipped
    assert np.allclose(np.linalg.norm(A2[1][:, [0, 2]], axis=0), 1.0)  # others rescaled


# -------------------------------------------------------------------- cost --

def test_cost_forms_agree_and_vanish_on_exact_data():
    rng = np.random.default_rng(6)
    dims = [8, 9, 10]
    A = [rng.random((I, 4)) for I in dims]
    X = o.cp_full(A)
    assert o.cost_explicit(X, A) < 1e-10 * np.linalg.norm(X)
    B = [a + 0.1 * rng.random(a.shape) for a in A]
    M = o.mttkrp(X, B, 2)
    V = o.partial_cwise([b.T @ b for b in B], 2)
    fe = o.cost_explicit(X, B)
    fg = o.cost_gram(o.square_norm(X), M, B[2], V, B[2].T @ B[2])
    ref = np.linalg.norm(X - o.cp_full(B))
    assert fe == pytest.approx(ref, rel=1e-12)
    assert fg == pytest.approx(ref, rel=1e-9)


# --------------------------------------------------------------------- ALS --

def _c1_like(dims=(30, 31, 32), R=4, snr=30.0):
    X, At = o.make_tensor(list(dims), R, seed=0, snr_db=snr)
    init = o.gen_factors(list(dims), R, seed=1)
    return X, At, init


def test_als_noiseless_recovery():
    X, At = o.make_tensor([20, 21, 22], 3, seed=0)
    init = o.gen_factors([20, 21, 22], 3, seed=1)
    # the default absolute NNLS tolerances (1e-2) stall near 3e-4 on this scale;
    # tight tolerances let ALS reach the exact low-rank fit
    F, st, log = o.cpd(X, 3, init, max_iter=200, threshold_error=1e-6, delta1=1e-12, delta2=1e-12)
    assert st["rel_cost"] < 1e-6 and st["ao_iter"] < 200
    assert st["rel_cost"] == pytest.approx(np.linalg.norm(X - o.cp_full(F)) / np.linalg.norm(X), rel=1e-6)


def test_als_status_fields_and_noise_floor():
    X, At, init = _c1_like()
    F, st, log = o.cpd(X, 4, init, max_iter=50, threshold_error=0.0)
    assert st["ao_iter"] == 50 and len(log) == 50
    assert st["frob_sq"] == pytest.approx(np.sum(X ** 2), rel=1e-13)           # ||X||^2, squared
    assert st["f_value"] == pytest.approx(np.linalg.norm(X - o.cp_full(F)), rel=1e-9)
    floor = 10 ** (-30 / 20)
    assert 0.9 * floor < st["rel_cost"] < 1.05 * floor
    assert all(f >= 0 for f in F[0].ravel())


@pytest.mark.parametrize("form,cost", [(o.MTTKRP_STANDARD, o.COST_GRAM), (o.MTTKRP_DIMTREE, o.COST_GRAM),
                                       (o.MTTKRP_BLOCKED, o.COST_EXPLICIT)])
def test_als_forms_agree(form, cost):
    X, At, init = _c1_like()
    F0, st0, log0 = o.cpd(X, 4, init, max_iter=25, threshold_error=0.0)
    F1, st1, log1 = o.cpd(X, 4, init, max_iter=25, threshold_error=0.0, mttkrp_form=form, cost_form=cost)
    assert [l["nnls_iters"] for l in log0] == [l["nnls_iters"] for l in log1]
    assert [l["accel_accepted"] for l in log0] == [l["accel_accepted"] for l in log1]
    for a, b in zip(F0, F1):
        assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max()


def test_threshold_termination():
    X, At, init = _c1_like(snr=40.0)
    F, st, log = o.cpd(X, 4, init, max_iter=50, threshold_error=0.05)
    assert 1 < st["ao_iter"] < 50 and st["rel_cost"] < 0.05
    assert all(rec["rel_cost"] >= 0.05 for rec in log[:-1])      # stops at the first sweep below


def test_accel_escalation_bookkeeping():
    X, At, init = _c1_like()
    F, st, log = o.cpd(X, 4, init, max_iter=40, threshold_error=0.0)
    coeff, fail = 3, 0
    for rec in log:
        if rec["accel_tried"]:
            if not rec["accel_accepted"]:
                fail += 1
            if fail == 5:
                fail, coeff = 0, coeff + 1
    assert (coeff, fail) == (st["accel_coeff"], st["accel_fail"])
    assert log[0]["accel_tried"] == 0 and all(r["accel_tried"] for r in log[1:-1])


def test_zero_column_fallback():
    """An MTTKRP that is <= 0 everywhere is driven to zero by NNLS, which
    triggers the 0.9 new + 0.1 old blend (Cpd.hpp:206-207)."""
    rng = np.random.default_rng(12)
    dims = [6, 7, 8]
    X = -o.cp_full([rng.random((I, 2)) for I in dims])      # negative data: M_0 <= 0
    init = [rng.random((I, 3)) for I in dims]
    F, st, log = o.cpd(X, 3, init, max_iter=1, threshold_error=0.0, normalization=0, acceleration=0)
    assert log[0]["zero_col_fallback"][0] == 1
    assert np.allclose(F[0], 0.1 * init[0], rtol=1e-14, atol=0)   # new factor is all zeros


def test_unconstrained_and_orthogonal_branches():
    rng = np.random.default_rng(13)
    dims = [10, 11, 12]
    X = o.cp_full([rng.random((I, 3)) for I in dims])
    init = [rng.random((I, 3)) for I in dims]
    cons = [o.NONNEGATIVITY, o.UNCONSTRAINED, o.ORTHOGONALITY]
    F, st, log = o.cpd(X, 3, init, max_iter=10, threshold_error=0.0, constraints=cons)
    assert np.allclose(F[2].T @ F[2], np.eye(3), atol=1e-10)   # orthogonal columns (last factor untouched by Normalize)
    assert np.isfinite(st["f_value"])


@pytest.mark.parametrize("n,kind", [(1, "scalar"), (2, "pair"), (7, "indefinite"), (33, "indefinite"),
                                    (64, "psd_rankdef"), (9, "diag")])
def test_sym_extremes_edge_cases(n, kind):
    """Singular-value extremes = max/min |eigenvalue|: indefinite spectra exercise the
    bracketing of the eigenvalues around 0, a diagonal input the no-reflection branch."""
    rng = np.random.default_rng(n)
    if kind in ("scalar", "pair", "indefinite"):
        B = rng.standard_normal((n, n))
        V = 0.5 * (B + B.T)
    elif kind == "psd_rankdef":
        B = rng.random((n, n // 2))
        V = B @ B.T                                # rank n/2: smallest |lambda| ~ rounding level
    else:
        V = np.diag(rng.standard_normal(n))
    w = np.abs(np.linalg.eigvalsh(V))
    L, mu, _ = o.sym_extremes(V)
    tol = 4 * n * 2.2e-16 * w.max()
    assert abs(L - w.max()) <= tol
    assert abs(mu - w.min()) <= tol


# ------------------------------------------- line search, Normalize, update --

def test_line_search_step_and_accelerated_cost():
    """line_search_accel (Cpd.hpp:247-278, cost :162-169): at sweep k the candidate is
    A_hat = A_old + s (A - A_old) with s = (k + 1)^(1 / accel_coeff), A the sweep's
    post-Normalize factors and A_old the previous sweep's post-Normalize factors taken
    BEFORE that sweep's own line search (Cpd.hpp:346-353).  A run stopped at sweep k
    returns exactly those factors (termination precedes the line search), so f_accel of a
    longer run is checked against the explicit residual ||X - [[A_hat]]||_F computed here,
    and the accept rule against f_value > f_accel."""
    X, _, init = _c1_like()
    K = 40
    _, stK, logK = o.cpd(X, 4, init, max_iter=K, threshold_error=0.0)
    assert stK["accel_coeff"] > 3                       # the coefficient escalates inside the run
    prev = o.cpd(X, 4, init, max_iter=1, threshold_error=0.0)
    after_accept = after_escalation = 0
    for k in range(2, K):
        cur = o.cpd(X, 4, init, max_iter=k, threshold_error=0.0)
        A_old, A, coeff = prev[0], cur[0], cur[1]["accel_coeff"]
        s = (k + 1) ** (1.0 / coeff)
        A_hat = [a0 + s * (a - a0) for a0, a in zip(A_old, A)]
        f_hat = np.linalg.norm(X - o.cp_full(A_hat))
        rec = logK[k - 1]
        assert rec["accel_tried"] == 1
        assert rec["f_accel"] == pytest.approx(f_hat, rel=1e-11), k
        assert rec["f_value"] == pytest.approx(cur[1]["f_value"], rel=1e-14)
        assert rec["accel_accepted"] == int(rec["f_value"] > rec["f_accel"])
        after_accept += logK[k - 2]["accel_accepted"]
        after_escalation += coeff > 3
        prev = cur
    assert after_accept > 0 and after_escalation > 0      # both cases a wrong A_old / s would miss


@pytest.mark.parametrize("N,wf", [(3, 1), (4, 1), (4, 2), (3, 2), (4, 3)])
def test_normalize_weight_factor(N, wf):
    """Normalize with the weights loaded on factor wf != 0 (Normalize.hpp:53-67, 92-174):
    factors j in [0, N-2] other than wf get unit columns, factor wf takes the product of
    their norms, the last factor is untouched unless it is wf, [[A]] and the Gram
    identities hold afterwards."""
    rng = np.random.default_rng(10 * N + wf)
    dims = [5, 6, 7, 4][:N]
    A = [rng.random((I, 3)) + 0.1 for I in dims]
    G = [a.T @ a for a in A]
    X0 = o.cp_full(A)
    G2, A2 = o.normalize(G, A, weight_factor=wf)
    assert np.allclose(o.cp_full(A2), X0, rtol=1e-13, atol=0)
    scaled = [j for j in range(N - 1) if j != wf]
    w = np.ones(3)
    for j in scaled:
        nj = np.linalg.norm(A[j], axis=0)
        w *= nj
        assert np.allclose(A2[j], A[j] / nj, rtol=1e-14, atol=0)
    assert np.allclose(A2[wf], A[wf] * w, rtol=1e-14, atol=0)
    if wf != N - 1:
        assert np.array_equal(A2[-1], A[-1])
    for j in range(N):
        assert np.allclose(G2[j], A2[j].T @ A2[j], rtol=1e-12, atol=0)


def test_normalize_weight_factor_zero_test():
    """With wf != 0 the zero test still reads the last entry of the norm list
    (Normalize.hpp:131-132): the last factor's norm when wf < N-1, factor N-2's when wf is
    the last factor."""
    rng = np.random.default_rng(4)
    A = [rng.random((4, 3)) + 0.1 for _ in range(3)]
    A[2][:, 1] = 0.0
    G = [a.T @ a for a in A]
    _, A2 = o.normalize(G, A, weight_factor=1)
    assert np.array_equal(A2[0][:, 1], A[0][:, 1])                     # column 1 skipped
    assert np.allclose(np.linalg.norm(A2[0][:, [0, 2]], axis=0), 1.0)
    B = [a.copy() for a in A]
    B[1][:, 2] = 0.0                                                   # factor N-2 zero, wf = last
    _, B2 = o.normalize([b.T @ b for b in B], B, weight_factor=2)
    assert np.array_equal(B2[0][:, 2], B[0][:, 2])
    assert np.allclose(np.linalg.norm(B2[0][:, [0]], axis=0), 1.0)


def _update_problem(I=40, R=6, seed=0):
    rng = np.random.default_rng(seed)
    B = rng.random((3 * R, R))
    V = B.T @ B + 0.5 * np.eye(R)
    M = rng.standard_normal((I, R)) + 1.0
    A0 = rng.random((I, R))
    return V, M, A0


def test_update_factor_unconstrained_solves_normal_equations():
    """Unconstrained branch A = M V^-1 (Cpd.hpp:196-200): A V = M to rounding, and the
    returned Gram is A^T A."""
    V, M, A0 = _update_problem()
    A, G, it, fb = o.update_factor(o.UNCONSTRAINED, V, M, A0)
    assert it == 0 and fb == 0
    assert np.linalg.norm(A @ V - M) <= 1e-12 * np.linalg.norm(M)
    assert np.allclose(G, A.T @ A, rtol=1e-13, atol=0)


def test_update_factor_orthogonality_is_polar_factor():
    """Orthogonality branch A = M (M^T M)^-1/2 (Cpd.hpp:210-219) is the orthonormal polar
    factor U W^T of M = U S W^T (SVD)."""
    V, M, A0 = _update_problem(seed=1)
    A, G, _, _ = o.update_factor(o.ORTHOGONALITY, V, M, A0)
    U, _, Wt = np.linalg.svd(M, full_matrices=False)
    assert np.abs(A - U @ Wt).max() <= 1e-11
    assert np.allclose(G, np.eye(M.shape[1]), atol=1e-12)


def test_update_factor_nonnegative_and_passive_branches():
    """Nonnegativity = the Nesterov solve from the previous factor (NesterovMNLS.hpp:181-234),
    blended 0.9 / 0.1 with the previous factor when a column vanishes (Cpd.hpp:206-207);
    sparsity and constant leave the factor as it is; every branch returns A^T A."""
    V, M, A0 = _update_problem(seed=2)
    A, G, it, fb = o.update_factor(o.NONNEGATIVITY, V, M, A0)
    An, itn, _, _ = o.nesterov(V, M, A0, 1e-2, 1e-2, 50)
    assert it == itn and fb == 0 and np.array_equal(A, An)
    assert np.allclose(G, A.T @ A, rtol=1e-13, atol=0)
    A, G, it, fb = o.update_factor(o.NONNEGATIVITY, V, -np.abs(M), A0)   # M <= 0: a column vanishes
    An, _, _, _ = o.nesterov(V, -np.abs(M), A0, 1e-2, 1e-2, 50)
    assert np.abs(An).sum(axis=0).min() == 0.0
    assert fb == 1 and np.array_equal(A, 0.9 * An + 0.1 * A0)
    for c in (o.SPARSITY, o.CONSTANT):
        A, G, it, fb = o.update_factor(c, V, M, A0)
        assert np.array_equal(A, A0) and np.allclose(G, A0.T @ A0, rtol=1e-13, atol=0)


def test_forced_line_search_outcomes():
    """The test hook that makes the oracle follow given line-search outcomes: -1 keeps the
    oracle's own test, a forced outcome that differs from it is flagged, and the counters of
    Cpd.hpp:270-277 follow the forced outcomes."""
    X, _, init = _c1_like()
    F, st, log = o.cpd(X, 4, init, max_iter=12, threshold_error=0.0)
    own = [l["accel_accepted"] for l in log]
    F1, st1, log1 = o.cpd(X, 4, init, max_iter=12, threshold_error=0.0, force_accept=[-1] * 12)
    assert [l["accel_accepted"] for l in log1] == own and not any(l["accel_overridden"] for l in log1)
    for a, b in zip(F, F1):
        assert np.array_equal(a, b)
    k = next(i for i in range(1, 11) if log[i]["accel_tried"])
    force = [-1] * 12
    force[k] = 1 - own[k]
    _, st2, log2 = o.cpd(X, 4, init, max_iter=12, threshold_error=0.0, force_accept=force)
    assert log2[k]["accel_overridden"] == 1 and log2[k]["accel_accepted"] == 1 - own[k]
    assert [l["accel_accepted"] for l in log2[:k]] == own[:k]
    coeff, fail = 3, 0
    for rec in log2:
        if rec["accel_tried"] and not rec["accel_accepted"]:
            fail += 1
            if fail == 5:
                fail, coeff = 0, coeff + 1
    assert (coeff, fail) == (st2["accel_coeff"], st2["accel_fail"])
