"""GPU parity: the HIP path through the C ABI against the FP64 CPU oracle on the
same inputs.  Integer-valued golden data must match exactly; FP64 runs within
the tolerances written in each test (1e-12 relative for the plain FP64 path's
factors and costs).  ALS runs compare every discrete decision; a line-search test
whose two costs tie to rounding may go either way, so the oracle is run on the
device's line-search outcomes and every sweep where they differ from its own test
must be such a tie -- the final factors are compared in every case.  Full-size
checks use the closed form of an exact low-rank MTTKRP (800^3, 200^4) and a
three-sweep 800^3 ALS against the oracle."""
import os
import sys

import numpy as np
import pytest

import oracle as o
import parity_util as U
import ptn_hip as P

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module")
def h():
    say("create handle")
    with P.Handle(0) as hd:
        yield hd


# ------------------------------------------------------------------ MTTKRP --

@pytest.mark.parametrize("name", ["mttkrp_3x4x5_r2.npz", "mttkrp_2x3x4x5_r3.npz"])
def test_mttkrp_golden_exact(h, golden_dir, name):
    say("golden", name)
    g = np.load(os.path.join(golden_dir, name))
    dims = [int(d) for d in g["dims"]]
    X = g["X"].reshape(dims, order="F")
    A = [g["A%d" % k] for k in range(len(dims))]
    h.set_tensor(X)
    for mode in range(len(dims)):
        assert np.array_equal(h.mttkrp(A, mode), g["M%d" % mode]), mode
    if len(dims) in (3, 4):
        D = h.mttkrp_dimtree(A)
        for mode in range(len(dims)):
            assert np.array_equal(D[mode], g["M%d" % mode]), mode


@pytest.mark.parametrize("dims,R", [([100, 100, 100], 10), ([37, 23, 41], 5), ([17, 65, 33], 33),
                                    ([64, 16, 48], 64), ([9, 11, 13, 7], 17), ([3, 5, 4, 2, 3], 6),
                                    ([20, 20, 30, 31], 40)])
def test_mttkrp_vs_oracle(h, dims, R):
    """Every mode and form against the oracle; [20, 20, 30, 31] gives the order-4 left node
    three k splits (400 rows, k = 930) whose last ends in a partial 16-k step."""
    say("mttkrp", dims, R)
    rng = np.random.default_rng(sum(dims) + R)
    X = np.asfortranarray(rng.standard_normal(dims))
    A = [rng.random((I, R)) for I in dims]
    h.set_tensor(X, storage=P.TENSOR_BOTH)
    for mode in range(len(dims)):
        ref = o.mttkrp(X, A, mode)
        got = h.mttkrp(A, mode)
        assert P.rel_err(got, ref) < 1e-13, (mode, P.rel_err(got, ref))
        # FP32 tensor stream with FP64 matrix cores: bounded by the FP32 rounding of X
        got32 = h.mttkrp(A, mode, arith=P.ARITH_F32_F64)
        assert P.rel_err(got32, ref) < 1e-6
        # FP32 matrix cores (FP32 products, FP32 accumulation) and bf16x3 split products
        assert P.rel_err(h.mttkrp(A, mode, arith=P.ARITH_F32), ref) < 2e-5, mode
        assert P.rel_err(h.mttkrp(A, mode, arith=P.ARITH_BF16X3), ref) < 2e-4, mode
    if len(dims) in (3, 4):
        D = h.mttkrp_dimtree(A)
        D3 = h.mttkrp_dimtree(A, arith=P.ARITH_BF16X3)
        D2 = h.mttkrp_dimtree(A, arith=P.ARITH_F32)
        for mode in range(len(dims)):
            ref = o.mttkrp(X, A, mode)
            assert P.rel_err(D[mode], ref) < 1e-13, mode
            assert P.rel_err(D2[mode], ref) < 2e-5, mode
            assert P.rel_err(D3[mode], ref) < 2e-4, mode


def test_mttkrp_full_size_closed_form(h):
    """800^3 exact rank-32 tensor: M_n = B_n (Hadamard_{k!=n} B_k^T A_k)."""
    dims, Rt, R = [800, 800, 800], 32, 32
    say("full size closed form", dims)
    h.generate_tensor(dims, Rt, seed=0, snr_db=None, storage=P.TENSOR_BOTH)
    B = o.gen_factors(dims, Rt, seed=0)
    A = o.gen_factors(dims, R, seed=1)
    for mode in range(3):
        H = np.ones((Rt, R))
        for k in range(3):
            if k != mode:
                H *= B[k].T @ A[k]
        ref = B[mode] @ H
        assert P.rel_err(h.mttkrp(A, mode), ref) < 1e-12, mode
    D = h.mttkrp_dimtree(A)
    D3 = h.mttkrp_dimtree(A, arith=P.ARITH_BF16X3)
    D2 = h.mttkrp_dimtree(A, arith=P.ARITH_F32)
    for mode in range(3):
        H = np.ones((Rt, R))
        for k in range(3):
            if k != mode:
                H *= B[k].T @ A[k]
        ref = B[mode] @ H
        assert P.rel_err(D[mode], ref) < 1e-12, mode
        say("full size", mode, "fp32 mfma err", P.rel_err(D2[mode], ref), "bf16x3 err", P.rel_err(D3[mode], ref))
        assert P.rel_err(D2[mode], ref) < 1e-5, mode
        assert P.rel_err(D3[mode], ref) < 1e-4, mode


def test_left_node_over_4gb_k_range_closed_form(h):
    """1100 x 1000 x 1000 FP32 (4.4 GB): the left node's k range spans more than 4 GB, so the
    strided contraction loads through 64-bit addresses instead of a raw buffer; the tree's
    MTTKRPs of the exact rank-8 tensor still match the closed form M_n = B_n (o B_k^T A_k)."""
    dims, Rt, R = [1100, 1000, 1000], 8, 32
    say("left node over a 4 GB k range", dims)
    h.generate_tensor(dims, Rt, seed=0, snr_db=None, storage=P.TENSOR_F32)
    B = o.gen_factors(dims, Rt, seed=0)
    A = o.gen_factors(dims, R, seed=1)
    for ar, tol in ((P.ARITH_F32, 1e-5), (P.ARITH_BF16X3, 1e-4)):
        D = h.mttkrp_dimtree(A, arith=ar)
        for mode in range(3):
            H = np.ones((Rt, R))
            for k in range(3):
                if k != mode:
                    H *= B[k].T @ A[k]
            err = P.rel_err(D[mode], B[mode] @ H)
            say("arith", ar, "mode", mode, "rel err", err)
            assert err < tol, (ar, mode, err)
    h.set_tensor(np.ones((2, 2, 2), order="F"))          # release the 4.4 GB


# --------------------------------------------------------- generator / cost --

def test_generator_matches_oracle(h):
    say("generator")
    dims = [30, 20, 10]
    A_gpu = h.generate_factors(dims, 7, seed=5)
    A_cpu = o.gen_factors(dims, 7, seed=5)
    for a, b in zip(A_gpu, A_cpu):
        assert np.array_equal(a, b)                       # integer arithmetic + exact conversion
    h.generate_tensor(dims, 4, seed=3, snr_db=None)
    X0, _ = o.make_tensor(dims, 4, seed=3)
    assert h.norm_sq() == pytest.approx(np.sum(X0 ** 2), rel=1e-14)
    h.generate_tensor(dims, 4, seed=3, snr_db=20.0)
    Xn, _ = o.make_tensor(dims, 4, seed=3, snr_db=20.0)
    assert h.norm_sq() == pytest.approx(np.sum(Xn ** 2), rel=1e-12)


def test_cost_explicit_matches_oracle(h):
    say("explicit cost")
    X, At = o.make_tensor([40, 30, 20], 5, seed=0, snr_db=25.0)
    init = o.gen_factors([40, 30, 20], 5, seed=1)
    h.set_tensor(X)
    assert h.cost_explicit(init) == pytest.approx(o.cost_explicit(X, init), rel=1e-13)


# ------------------------------------------------------------ R x R pieces --

def test_sym_extremes_match_oracle(h, golden_dir):
    g = np.load(os.path.join(golden_dir, "eig_extremes.npz"))
    for i in range(int(g["count"])):
        V = g["V%d" % i]
        say("eig", V.shape)
        L, mu = h.sym_extremes(V)
        Lo, muo, _ = o.sym_extremes(V)
        assert L == Lo and mu == muo                       # same arithmetic sequence: bitwise


@pytest.mark.parametrize("n,kind", [(1, "scalar"), (2, "pair"), (7, "indefinite"), (17, "indefinite"),
                                    (33, "indefinite"), (64, "indefinite"), (20, "negdef"), (64, "psd_rankdef"),
                                    (9, "diag"), (40, "blockdiag")])
def test_sym_extremes_edge_cases_match_oracle(h, n, kind):
    """Spectra the Nesterov constants rarely see: eigenvalues of both signs (the pair
    bracketed around 0), all negative, rank-deficient, and matrices with zero columns to
    eliminate (the no-reflection branch), at every padded size (16, 32, 64): bitwise."""
    say("eig edge", n, kind)
    rng = np.random.default_rng(100 + n)
    if kind in ("scalar", "pair", "indefinite"):
        B = rng.standard_normal((n, n))
        V = 0.5 * (B + B.T)
    elif kind == "negdef":
        B = rng.random((n, 2 * n))
        V = -(B @ B.T)
    elif kind == "psd_rankdef":
        B = rng.random((n, n // 2))
        V = B @ B.T
    elif kind == "diag":
        V = np.diag(rng.standard_normal(n))
    else:                                                  # two blocks: zero entries below the first
        V = np.zeros((n, n))
        for a, b in ((0, n // 2), (n // 2, n)):
            B = rng.standard_normal((b - a, b - a))
            V[a:b, a:b] = 0.5 * (B + B.T)
    L, mu = h.sym_extremes(V)
    Lo, muo, _ = o.sym_extremes(V)
    assert L == Lo and mu == muo


@pytest.mark.parametrize("I,R,d", [(7, 4, 1e-2), (300, 32, 1e-2), (129, 64, 1e-2), (50, 10, 0.0), (20, 5, 1e30)])
def test_nesterov_matches_oracle(h, I, R, d):
    say("nnls", I, R, d)
    rng = np.random.default_rng(I + R)
    B = rng.random((3 * R, R))
    Z = B.T @ B
    W = rng.standard_normal((I, R)) * 4 + 2
    A0 = rng.random((I, R))
    Ag, itg, Lg, mug = h.nesterov(Z, W, A0, d, d, 50)
    Ao, ito, Lo, muo, = o.nesterov(Z, W, A0, d, d, 50)
    assert itg == ito
    assert (Lg, mug) == (Lo, muo)
    assert np.array_equal(Ag, Ao)                          # row arithmetic mirrors the oracle


@pytest.mark.parametrize("wf", [0, 1, 2, 3])
def test_normalize_matches_oracle(h, wf):
    """Normalize with the weights loaded on factor wf (Normalize.hpp:92-174): bitwise the
    oracle's, including the zero-norm branch, for every weight factor of orders 3 and 4."""
    rng = np.random.default_rng(3 + wf)
    for N in (3, 4):
        if wf >= N:
            continue
        dims = [7, 6, 5, 4][:N]
        A = [rng.random((I, 5)) for I in dims]
        A[-1][:, 2] = 0.0                                  # zero-norm branch on one column
        A[N - 2][:, 4] = 0.0                               # ... the one wf = N - 1 reads
        G = [a.T @ a for a in A]
        say("normalize N=%d wf=%d" % (N, wf))
        Gg, Ag = h.normalize(G, A, wf)
        Go, Ao = o.normalize(G, A, wf)
        for a, b in zip(Ag, Ao):           # a zero norm the test does not read divides by 0: NaN on both
            assert np.array_equal(a, b, equal_nan=True)
        for a, b in zip(Gg, Go):
            assert np.array_equal(a, b, equal_nan=True)


def test_order4_gemm_node_full_size_closed_form(h):
    """200^4 (6.4 GB FP32, 12.8 GB FP64), noiseless and of exact rank 64 with Exponential(1)
    factors: every MTTKRP of the tree -- the order-4 GEMM node in FP32 MFMA and bf16x3, and
    the FP64 path on the FP64 copy -- against the closed form M_n = B_n (o_{k!=n} B_k^T A_k).
    Tolerances: FP64 1e-12; FP32 rounding of X plus FP32 accumulation over k = 40000 in
    split chunks 1e-5 (FP32 MFMA) and 1e-4 (bf16x3)."""
    dims, Rt, R = [200, 200, 200, 200], 64, 64
    say("order-4 node 200^4 R=64 closed form")
    h.generate_tensor(dims, Rt, seed=2, snr_db=None, dist=1, storage=P.TENSOR_BOTH)
    B = o.gen_factors(dims, Rt, seed=2, dist=1)
    A = h.generate_factors(dims, R, seed=1)
    refs = []
    for mode in range(4):
        H = np.ones((Rt, R))
        for k in range(4):
            if k != mode:
                H *= B[k].T @ A[k]
        refs.append(B[mode] @ H)
    for ar, tol in ((P.ARITH_F64, 1e-12), (P.ARITH_F32, 1e-5), (P.ARITH_BF16X3, 1e-4)):
        M = h.mttkrp_dimtree(A, arith=ar)
        errs = [P.rel_err(M[k], refs[k]) for k in range(4)]
        say("arith", ar, "rel errs vs closed form", ["%.2e" % e for e in errs])
        assert max(errs) < tol, (ar, errs)
    h.set_tensor(np.ones((2, 2, 2), order="F"))          # release the 19 GB


# --------------------------------------------------------------------- ALS --

TOL = 1e-12       # plain FP64 path: factors, relative
COST_TOL = 2e-13  # |f_device^2 - f_oracle^2| / ||X||^2 (see _compare_runs)
TIE = 1e-12       # |f - f_accel| / f below which a line-search outcome rests on rounding


def oracle_on_device_path(X, R, init, logg, tie=TIE, **kw):
    return U.oracle_on_device_path(X, R, init, logg, tie, **kw)


def _compare_runs(Fg, stg, logg, Fo, sto, logo, tol=TOL, cost_tol=COST_TOL):
    """Every discrete decision equal; factors within tol.  Costs are compared on the scale
    of ||X||^2: the Gram-form cost f^2 = ||X||^2 - 2<M, A> + <V, G> is a difference of
    three terms of that size, each a sum of up to 10^6 products taken in another order on
    each side (the oracle sums ||X||^2 sequentially in chunks of 65 536, ~sqrt(65536) eps =
    3e-14 of rounding), so f^2 agrees to COST_TOL of ||X||^2 (measured <= 2.2e-14 here,
    6.5e-14 in the batch tests), which near the noise floor is ~1e-12 of f itself."""
    assert stg["ao_iter"] == sto["ao_iter"]
    assert [l["nnls_iters"] for l in logg] == [l["nnls_iters"] for l in logo]
    assert [l["accel_tried"] for l in logg] == [l["accel_tried"] for l in logo]
    assert [l["accel_accepted"] for l in logg] == [l["accel_accepted"] for l in logo]
    assert [l["zero_col_fallback"] for l in logg] == [l["zero_col_fallback"] for l in logo]
    frob = sto["frob_sq"]
    fgap = max(abs(a["f_value"] ** 2 - b["f_value"] ** 2) for a, b in zip(logg, logo)) / frob
    agap = max([abs(a["f_accel"] ** 2 - b["f_accel"] ** 2) for a, b in zip(logg, logo) if b["accel_tried"]]
               or [0.0]) / frob
    ferr = max(P.rel_err(a, b) for a, b in zip(Fg, Fo))
    say("vs oracle: final factors %.2e (tolerance %.1e); max |f^2 - f_o^2| / ||X||^2 %.2e, accelerated %.2e "
        "(tolerance %.1e)" % (ferr, tol, fgap, agap, cost_tol))
    assert fgap < cost_tol and agap < cost_tol
    assert (stg["accel_coeff"], stg["accel_fail"]) == (sto["accel_coeff"], sto["accel_fail"])
    assert ferr < tol


def _als_parity(h, X, R, init, oform=o.MTTKRP_STANDARD, **kw):
    """One device run through the C ABI against the oracle's run on its path.  The factor
    (cost) tolerance is 1e-12 (COST_TOL), or four times the spread between the oracle's own
    summation orders (per-mode, blocked and tree MTTKRP on the same path) where the problem
    amplifies rounding beyond that: the orthogonality branch's (M^T M)^-1/2 and the
    unconstrained M V^-1 reach a 2.5e-11 spread after ten sweeps in the oracle alone."""
    Fg, stg, logg = h.cpd(R, init, **kw)
    okw = {k: v for k, v in kw.items() if k != "mttkrp_form"}
    Fo, sto, logo = oracle_on_device_path(X, R, init, logg, mttkrp_form=oform, cost_form=o.COST_GRAM, **okw)
    spread, cspread = U.oracle_spread(X, R, init, logg, Fo, logo, oform, cost_form=o.COST_GRAM, **okw)
    say("oracle's own spread between summation orders: factors %.2e, |f^2 - f^2| / ||X||^2 %.2e"
        % (spread, cspread))
    _compare_runs(Fg, stg, logg, Fo, sto, logo, max(TOL, 4.0 * spread), max(COST_TOL, 4.0 * cspread))
    return Fg, stg, logg


def _oform(form):
    return o.MTTKRP_DIMTREE if form == P.MTTKRP_DIMTREE else o.MTTKRP_STANDARD


@pytest.mark.parametrize("form", [P.MTTKRP_STANDARD, P.MTTKRP_DIMTREE])
def test_als_c1_parity(h, form):
    """Config 1: 100^3, R=10, SNR 30 dB, init seed 1, 50 sweeps, reference defaults."""
    say("ALS C1 form", form)
    dims, R = [100, 100, 100], 10
    X, _ = o.make_tensor(dims, R, seed=0, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=1)
    h.set_tensor(X)
    Fg, stg, logg = _als_parity(h, X, R, init, mttkrp_form=form, oform=_oform(form), max_iter=50,
                                threshold_error=0.0)
    say("rel", stg["rel_cost"], "sweeps/s", stg["ao_iter"] / stg["sweep_seconds"])


@pytest.mark.parametrize("form", [P.MTTKRP_STANDARD, P.MTTKRP_DIMTREE])
def test_als_threshold_stop(h, form):
    """The relative-cost threshold ends the run at the same sweep as the oracle; the
    stopping sweep takes no acceleration step (termination precedes line_search_accel)."""
    dims, R = [60, 50, 40], 6
    X, _ = o.make_tensor(dims, R, seed=8, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=9)
    h.set_tensor(X)
    Fg, stg, logg = _als_parity(h, X, R, init, mttkrp_form=form, oform=_oform(form), max_iter=200,
                                threshold_error=0.05)
    say("threshold stop at", stg["ao_iter"])
    assert 1 < stg["ao_iter"] < 200
    assert logg[stg["ao_iter"] - 1]["accel_tried"] == 0


def test_als_c1_no_accel_no_norm(h):
    say("ALS C1 variant: acceleration and normalization off")
    dims, R = [100, 100, 100], 10
    X, _ = o.make_tensor(dims, R, seed=0, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=1)
    h.set_tensor(X)
    _als_parity(h, X, R, init, max_iter=50, threshold_error=0.0, acceleration=0, normalization=0)


@pytest.mark.parametrize("dims,R", [([12, 9, 10, 8], 4), ([30, 31, 32], 33)])
def test_als_small_parity(h, dims, R):
    say("ALS small", dims, R)
    X, _ = o.make_tensor(dims, min(R, 6), seed=2, snr_db=25.0)
    init = o.gen_factors(dims, R, seed=3)
    h.set_tensor(X)
    for form in (P.MTTKRP_STANDARD, P.MTTKRP_DIMTREE):
        _als_parity(h, X, R, init, mttkrp_form=form, oform=_oform(form), max_iter=15, threshold_error=0.0)


def test_als_c2_full_size_three_sweeps(h):
    """Config 2 at full size (800^3, true rank 32, SNR 20 dB, R = 32, init seed 1): three
    sweeps (the line search runs at sweep 2) of the plain FP64 path, per-mode and tree, against
    the oracle (OpenMP blocked MTTKRP, Gram-form cost) at 1e-12; the FP32-stream forms (tree,
    FP32 MFMA and bf16x3 with the fused line-search pass) end within 1e-6 of the oracle's
    relative error.  The device generates the tensor (bitwise the oracle's generator, tested
    above); its ||X||^2 is checked against the oracle's tensor."""
    dims, R, sweeps = [800, 800, 800], 32, 3
    say("C2 full size: oracle tensor")
    X, _ = o.make_tensor(dims, R, seed=0, snr_db=20.0)
    init = o.gen_factors(dims, R, seed=1)
    say("C2 full size: device tensor")
    h.generate_tensor(dims, R, seed=0, snr_db=20.0, storage=P.TENSOR_BOTH)
    assert h.norm_sq() == pytest.approx(float(np.sum(X * X)), rel=1e-12)
    kw = dict(max_iter=sweeps, threshold_error=0.0)
    say("C2 full size: oracle run")
    Fo, sto, logo = o.cpd(X, R, init, mttkrp_form=o.MTTKRP_BLOCKED, cost_form=o.COST_GRAM, threads=0, **kw)
    assert logo[1]["accel_tried"] == 1
    for form in (P.MTTKRP_STANDARD, P.MTTKRP_DIMTREE):
        say("C2 full size: FP64 form", form)
        Fg, stg, logg = h.cpd(R, init, mttkrp_form=form, **kw)
        assert [l["accel_accepted"] for l in logg] == [l["accel_accepted"] for l in logo]
        _compare_runs(Fg, stg, logg, Fo, sto, logo)
    for ar in (P.ARITH_F32, P.ARITH_BF16X3):
        Fc, stc, _ = h.cpd(R, init, mttkrp_form=P.MTTKRP_DIMTREE, arith=ar, fuse_accel=1, **kw)
        gap = abs(stc["rel_cost"] - sto["rel_cost"])
        ferr = max(P.rel_err(a, b) for a, b in zip(Fc, Fo))
        say("C2 full size: arith", ar, "rel error %.10f vs oracle %.10f (gap %.2e), factors %.2e"
            % (stc["rel_cost"], sto["rel_cost"], gap, ferr))
        assert gap < 1e-6
    h.set_tensor(np.ones((2, 2, 2), order="F"))          # release the 6 GB


@pytest.mark.parametrize("arith", [P.ARITH_F32, P.ARITH_BF16X3])
@pytest.mark.parametrize("dims,R", [([100, 100, 100], 10), ([24, 20, 18, 16], 7)])
def test_als_fp32_stream_and_fused_line_search(h, arith, dims, R):
    """FP32 tensor stream on the dimension tree, with and without the line-search pass
    fused into the next left node.  The paired pass computes the lower half of its
    output column by column exactly as the single left node does, so while the two runs
    take the same acceleration decisions they are bitwise equal apart from f_accel
    (whose fit term is summed in another order).  Against the FP64 oracle the final fit
    agrees to the FP32 stream's accuracy."""
    say("ALS fp32 stream", dims, R, "arith", arith)
    X, _ = o.make_tensor(dims, R, seed=0, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=1)
    h.set_tensor(X, storage=P.TENSOR_BOTH)
    kw = dict(max_iter=25, threshold_error=0.0, mttkrp_form=P.MTTKRP_DIMTREE, arith=arith)
    F0, st0, lg0 = h.cpd(R, init, fuse_accel=0, **kw)
    F1, st1, lg1 = h.cpd(R, init, fuse_accel=1, **kw)
    acc0 = [l["accel_accepted"] for l in lg0]
    acc1 = [l["accel_accepted"] for l in lg1]
    say("accept unfused", acc0, "fused", acc1)
    tried = [i for i, l in enumerate(lg0) if l["accel_tried"]]
    assert tried and [i for i, l in enumerate(lg1) if l["accel_tried"]] == tried
    # first line search: same state in both runs, the fit term <X, [[Ahat]]> summed in
    # another order; FP32 accumulation bounds its difference by ~1e-6 of ||X||^2
    frob = float(np.sum(X * X))
    i = tried[0]
    fa0, fa1 = lg0[i]["f_accel"], lg1[i]["f_accel"]
    say("first f_accel", fa0, fa1, "||X||^2", frob)
    assert abs(fa1 * fa1 - fa0 * fa0) < 1e-5 * frob
    if acc0 == acc1:
        assert [l["nnls_iters"] for l in lg0] == [l["nnls_iters"] for l in lg1]
        for a, b in zip(F0, F1):
            assert np.array_equal(a, b)
    # against the FP64 oracle on each run's own line-search path: FP32 errors scale with
    # ||X||^2 (the fit term <M, A> carries ~1e-7 of it), so squared residuals are compared on
    # that scale, and a line-search outcome is a tie when the oracle's two costs are within
    # the ~1e-4 relative error this puts on f (1e-3 allowed)
    rel_o = None
    for F, lg in ((F0, lg0), (F1, lg1)):
        Fo, sto, logo = oracle_on_device_path(X, R, init, lg, tie=1e-3, mttkrp_form=o.MTTKRP_DIMTREE,
                                              cost_form=o.COST_GRAM, max_iter=25, threshold_error=0.0)
        worst = max(abs(a["f_value"] ** 2 - b["f_value"] ** 2) for a, b in zip(lg, logo)) / frob
        rel = h.cost_explicit(F) / np.sqrt(frob)
        rel_o = o.cost_explicit(X, Fo) / np.sqrt(frob)
        say("on the device's path: max |f^2 - f_oracle^2| / ||X||^2 = %.2e; final rel cost %.10f, oracle %.10f"
            % (worst, rel, rel_o))
        assert worst < 1e-5
        assert abs(rel * rel - rel_o * rel_o) < 1e-5


def test_time_passes_events_leave_results_unchanged(h):
    """time_passes brackets the two tensor passes of every sweep with HIP events: the
    factors, costs and decisions are bitwise those of the untimed run, the order-3 tree
    reports both passes in every sweep that runs them (the left-node pass from sweep 1, the
    paired one from the line search on; the right-node pass in every sweep), and the
    per-sweep times stay within that sweep's wall time."""
    dims, R = [120, 110, 100], 12
    X, _ = o.make_tensor(dims, R, seed=4, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=5)
    h.set_tensor(X, storage=P.TENSOR_F32)
    kw = dict(max_iter=10, threshold_error=0.0, mttkrp_form=P.MTTKRP_DIMTREE, arith=P.ARITH_BF16X3, fuse_accel=1)
    F0, st0, lg0 = h.cpd(R, init, **kw)
    F1, st1, lg1 = h.cpd(R, init, time_passes=1, **kw)
    assert st0["f_value"] == st1["f_value"]
    assert [l["f_accel"] for l in lg0] == [l["f_accel"] for l in lg1]
    for a, b in zip(F0, F1):
        assert np.array_equal(a, b)
    assert all(l["t_pass_ms"] == [0.0, 0.0] for l in lg0)
    t_prev = 0.0
    for k, l in enumerate(lg1):
        left, right = l["t_pass_ms"]
        say("sweep", k + 1, "passes (ms)", left, right)
        assert right > 0.0
        if k + 1 < len(lg1):                 # the last sweep launches no further left-node pass
            assert left > 0.0
        assert left + right < 1e3 * (l["t_end"] - t_prev) + 1.0
        t_prev = l["t_end"]


@pytest.mark.parametrize("dims,R,seed,sweeps", [([5, 4, 6, 3, 4], 3, 5, 9), ([4, 3, 5, 3, 4, 2], 3, 6, 10),
                                                ([3, 4, 5, 3, 2, 3], 2, 7, 10)])
def test_als_higher_order_standard_form(h, dims, R, seed, sweeps):
    """Orders 5 and 6 through the per-mode form."""
    say("ALS order", len(dims), dims, R)
    X, _ = o.make_tensor(dims, R, seed=seed, snr_db=25.0)
    init = o.gen_factors(dims, R, seed=seed + 1)
    h.set_tensor(X)
    _als_parity(h, X, R, init, max_iter=sweeps, threshold_error=0.0)


def _fused_equals_unfused(h, R, init, kw):
    """The tree with FP32 contractions, with and without the line-search pass fused into
    the next left node: while the two runs take the same acceleration decisions they are
    bitwise equal apart from f_accel (its fit term is summed in another order)."""
    kt = dict(kw, mttkrp_form=P.MTTKRP_DIMTREE, arith=P.ARITH_F32)
    F0, st0, lg0 = h.cpd(R, init, fuse_accel=0, **kt)
    F1, st1, lg1 = h.cpd(R, init, fuse_accel=1, **kt)
    acc0 = [l["accel_accepted"] for l in lg0]
    acc1 = [l["accel_accepted"] for l in lg1]
    assert [l["accel_tried"] for l in lg0] == [l["accel_tried"] for l in lg1]
    say("fused line search: accept unfused", acc0, "fused", acc1)
    if acc0 == acc1:
        assert [l["nnls_iters"] for l in lg0] == [l["nnls_iters"] for l in lg1]
        assert st0["f_value"] == st1["f_value"]
        for a, b in zip(F0, F1):
            assert np.array_equal(a, b)
    else:
        k = next(i for i, (a, b) in enumerate(zip(acc0, acc1)) if a != b)
        assert abs(lg0[k]["f_value"] - lg0[k]["f_accel"]) <= 1e-6 * lg0[k]["f_value"], k


@pytest.mark.parametrize("cons", [[0, 0, 0], [2, 1, 1], [1, 0, 2], [1, 4, 1], [0, 2, 1, 3]])
def test_als_other_constraints(h, cons):
    """update_factor's other branches on the device (Cpd.hpp:189-228): unconstrained
    M V^-1 (LU with partial pivoting), orthogonality M (M^T M)^-1/2 (round-robin Jacobi),
    sparsity / constant (factor not updated), mixed with nonnegativity, both MTTKRP forms,
    against the oracle's update_factor; Normalize takes the first non-orthogonal factor."""
    dims = [30, 25, 20, 12][:len(cons)]
    R = 4
    say("constraints", cons)
    X, _ = o.make_tensor(dims, R, seed=13, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=14)
    h.set_tensor(X)
    kw = dict(max_iter=10, threshold_error=0.0, constraints=cons)
    for form, oform in ((P.MTTKRP_STANDARD, o.MTTKRP_STANDARD), (P.MTTKRP_DIMTREE, o.MTTKRP_DIMTREE)):
        Fg, stg, logg = _als_parity(h, X, R, init, mttkrp_form=form, oform=oform, **kw)
        for k, c in enumerate(cons):
            if c == 2:                                      # orthonormal columns
                assert np.allclose(Fg[k].T @ Fg[k], np.eye(R), atol=1e-10)
            if c == 4:                                      # constant: only rescaled (Normalize)
                ratio = Fg[k] / init[k]
                assert np.allclose(ratio, ratio[0:1, :], rtol=1e-12)


@pytest.mark.parametrize("dims,R", [([6, 5, 7, 4, 5], 3), ([4, 5, 3, 4, 3, 4], 3), ([3, 4, 3, 2, 3, 2, 3], 2),
                                    ([3, 2, 3, 2, 2, 3, 2, 3], 2)])
def test_dimtree_higher_orders(h, dims, R):
    """Orders 5-8 in the tree: the root's children (modes [0, ceil(N/2)) and the rest), each
    one pass over X; every MTTKRP against the oracle's, FP64 and FP32 stream; the fused
    line-search pass against the unfused one; then ALS in the tree form against the
    oracle's standard form."""
    say("tree order", len(dims), dims)
    rng = np.random.default_rng(len(dims))
    X = np.asfortranarray(rng.standard_normal(dims))
    A = [rng.random((I, R)) for I in dims]
    h.set_tensor(X, storage=P.TENSOR_BOTH)
    D = h.mttkrp_dimtree(A)
    D32 = h.mttkrp_dimtree(A, arith=P.ARITH_F32_F64)
    for mode in range(len(dims)):
        ref = o.mttkrp(X, A, mode)
        assert P.rel_err(D[mode], ref) < 1e-13, mode
        assert P.rel_err(D32[mode], ref) < 1e-6, mode
    Xc, _ = o.make_tensor(dims, R, seed=len(dims) + 20, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=len(dims) + 21)
    h.set_tensor(Xc, storage=P.TENSOR_BOTH)
    kw = dict(max_iter=8, threshold_error=0.0)
    _fused_equals_unfused(h, R, init, kw)
    _als_parity(h, Xc, R, init, mttkrp_form=P.MTTKRP_DIMTREE, **kw)


@pytest.mark.parametrize("dims", [[6, 7, 8], [70, 37, 45]])
def test_zero_column_fallback_on_gpu(h, dims):
    """A negative tensor drives every Nesterov solve to zero: each mode takes the
    0.9 A + 0.1 A_old fallback, and its Gram (one or several 32-row chunks) feeds the
    next modes, so every factor and the cost check the fallback's Gram too."""
    say("zero-column fallback", dims)
    rng = np.random.default_rng(12)
    X = -o.cp_full([rng.random((I, 2)) for I in dims])
    init = [rng.random((I, 3)) for I in dims]
    h.set_tensor(X)
    kw = dict(max_iter=2, threshold_error=0.0, normalization=0, acceleration=0)
    Fg, stg, logg = h.cpd(3, init, **kw)
    Fo, sto, logo = o.cpd(X, 3, init, cost_form=o.COST_GRAM, **kw)
    assert logg[0]["zero_col_fallback"][0] == 1
    assert [l["zero_col_fallback"] for l in logg] == [l["zero_col_fallback"] for l in logo]
    for a, b in zip(Fg, Fo):
        assert np.array_equal(a, b)
    assert stg["f_value"] == pytest.approx(sto["f_value"], rel=1e-12)


def test_invalid_arguments_raise(h):
    say("error paths")
    h.set_tensor(np.ones((4, 5, 6), order="F"))
    init = [np.ones((I, 3)) for I in (4, 5, 6)]
    with pytest.raises(P.PtnError):
        h.cpd(65, [np.ones((I, 65)) for I in (4, 5, 6)])          # rank above the wave-per-row limit
    with pytest.raises(P.PtnError):
        h.cpd(3, init, constraints=[1, 7, 1])                       # no such constraint
    with pytest.raises(P.PtnError):
        h.mttkrp(init, 5)                                           # bad mode
    with pytest.raises(P.PtnError):
        h.mttkrp(init, 0, arith=P.ARITH_F32_F64)                    # no FP32 copy uploaded
    h.set_tensor(np.ones((4, 5, 6), order="F"), storage=P.TENSOR_BOTH)
    with pytest.raises(P.PtnError):
        h.cpd(3, init, fuse_accel=1, arith=P.ARITH_F32)             # paired pass needs the tree
    with pytest.raises(P.PtnError):
        h.cpd(3, init, fuse_accel=1, mttkrp_form=P.MTTKRP_DIMTREE)  # ... and FP32-output contractions
