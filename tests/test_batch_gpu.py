"""Batched independent CPDs on the GPU against per-problem runs of the CPU oracle.

The batch has no reference counterpart (the reference runs one cpd() per tensor); it is
pinned by running the oracle's cpd() (standard MTTKRP, Gram-form cost) on every problem
with the same FP32-rounded tensor and the same initial factors.  The MTTKRP summation
order differs from the oracle's, so costs and factors agree to the tolerances written
below (1e-12 relative for the FP64 batch), while every discrete decision (Nesterov step
counts, accepted line searches, threshold stops) must agree exactly.  A line search whose
two costs tie to rounding may go either way: the oracle follows the device's outcomes and
every sweep where they differ from its own test must be such a tie."""
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


# mode sizes on both sides of the lane-mapping branches (I0 < 64, = 64, > 64, = 128),
# ranks 1..16 across the four rank paddings, a few tiny modes
EDGE = [((128, 128, 128), 16), ((64, 65, 63), 13), ((16, 17, 18), 4), ((100, 37, 90), 9), ((2, 3, 4), 2),
        ((33, 128, 16), 16), ((127, 5, 64), 7), ((20, 20, 20), 1), ((48, 80, 112), 12), ((65, 64, 63), 8),
        ((17, 120, 33), 5), ((90, 16, 128), 11)]


def _problems(n_random, seed):
    rng = np.random.default_rng(seed)
    probs = list(EDGE)
    for _ in range(n_random):
        probs.append((tuple(int(v) for v in rng.integers(2, 129, 3)), int(rng.integers(1, 17))))
    return probs


def _tensors(probs, snr_db=25.0, seed0=0):
    """FP64 tensors from the oracle's generator, rounded to FP32 (what the device stores)."""
    Xs = []
    for p, (dims, R) in enumerate(probs):
        X, _ = o.make_tensor(list(dims), R, seed=seed0 + p, snr_db=snr_db)
        Xs.append(np.asfortranarray(X.astype(np.float32).astype(np.float64)))
    return Xs


TOL = 1e-12       # FP64 batch: factors against the oracle, relative
COST_TOL = 2e-13  # |f_device^2 - f_oracle^2| / ||X||^2: the Gram-form cost is a difference of
                  # three terms of size ||X||^2, each a sum of up to 2*10^6 products taken in
                  # another order on each side (measured <= 6.5e-14)
TIE = 1e-12       # |f - f_accel| / f below which a line-search outcome rests on rounding


def oracle_problem(X, R, init, logg, **kw):
    """The oracle's cpd() of one problem on the device's line-search outcomes (every sweep
    where they differ from the oracle's own test must be a tie to TIE), and the factor
    tolerance: 1e-12, or four times the oracle's own spread between its summation orders
    (per-mode, blocked and tree MTTKRP on the same path) where the problem amplifies
    rounding beyond that (up to 1.6e-12 here)."""
    Fo, sto, logo = U.oracle_on_device_path(X, R, init, logg, TIE, cost_form=o.COST_GRAM, **kw)
    spread, _ = U.oracle_spread(X, R, init, logg, Fo, logo, o.MTTKRP_STANDARD, cost_form=o.COST_GRAM, **kw)
    return Fo, sto, logo, max(TOL, 4.0 * spread)


def _check_problem(p, Fg, stg, logg, Fo, sto, logo, tol=TOL):
    """Sweep by sweep: equal step counts and line-search outcomes, costs within COST_TOL of
    ||X||^2; final factors within tol, status and counters.  Returns (factor gap, cost gap)."""
    assert stg["ao_iter"] == sto["ao_iter"], p
    frob = sto["frob_sq"]
    cgap = 0.0
    for k, (a, b) in enumerate(zip(logg, logo)):
        cgap = max(cgap, abs(a["f_value"] ** 2 - b["f_value"] ** 2) / frob)
        assert a["nnls_iters"] == b["nnls_iters"][:3], (p, k)
        assert a["accel_tried"] == b["accel_tried"], (p, k)
        assert a["accel_accepted"] == b["accel_accepted"], (p, k)
        if b["accel_tried"]:
            cgap = max(cgap, abs(a["f_accel"] ** 2 - b["f_accel"] ** 2) / frob)
    assert cgap < COST_TOL, (p, cgap)
    assert (stg["accel_coeff"], stg["accel_fail"]) == (sto["accel_coeff"], sto["accel_fail"]), p
    err = max(P.rel_err(a, b) for a, b in zip(Fg, Fo))
    assert err < tol, (p, err)
    return err, cgap


def test_batch_parity_vs_oracle(h):
    probs = _problems(12, seed=11)
    say("batch parity:", len(probs), "problems")
    Xs = _tensors(probs)
    init = [o.gen_factors(list(d), R, seed=100 + p) for p, (d, R) in enumerate(probs)]
    h.batch_set(Xs, [R for _, R in probs])
    kw = dict(max_iter=12, threshold_error=0.0)
    Fg, stg, logg, rep = h.cpd_batch(init, block=256, grid=40, **kw)
    say("report", rep)
    worst = [0.0, 0.0]
    for p, ((dims, R), X) in enumerate(zip(probs, Xs)):
        Fo, sto, logo, tol = oracle_problem(X, R, init[p], logg[p], **kw)
        if tol > TOL:
            say("problem", p, probs[p], "factor tolerance %.2e from the oracle's own spread" % tol)
        worst = [max(a, b) for a, b in zip(worst, _check_problem(p, Fg[p], stg[p], logg[p], Fo, sto, logo, tol))]
        assert stg[p]["frob_sq"] == pytest.approx(sto["frob_sq"], rel=1e-13)
    say("over the batch vs oracle: max factor rel diff %.2e, max |f^2 - f_o^2| / ||X||^2 %.2e" % tuple(worst))


def test_batch_split_problems_independent_of_blocks(h):
    """A problem with far more work than the batch's mean has its MTTKRP sums cut in pieces
    (8 here) that up to 8 blocks share, exchanging them at a group barrier: on 1, 3 and 40
    blocks (1, 3 and 8 members) the factors are bitwise equal, and the run matches the
    oracle like an unsplit problem (both arithmetics)."""
    probs = [((120, 110, 100), 12)] + [((6 + p % 5, 5 + p % 3, 7), 1 + p % 4) for p in range(40)]
    say("split batch:", probs[0], "+", len(probs) - 1, "small problems")
    Xs = _tensors(probs, seed0=500)
    init = [o.gen_factors(list(d), R, seed=600 + p) for p, (d, R) in enumerate(probs)]
    h.batch_set(Xs, [R for _, R in probs])
    kw = dict(max_iter=10, threshold_error=0.0)
    for arith in (P.ARITH_F64, P.ARITH_F32):
        runs = []
        for grid in (1, 3, 40):
            F, st, lg, rep = h.cpd_batch(init, block=512, grid=grid, arith=arith, **kw)
            say("arith", arith, "grid", grid, "report", rep)
            assert rep["split_problems"] == 1 and rep["max_pieces"] == 8
            runs.append((F, st, lg))
        for F, st, _ in runs[1:]:
            assert [s_["f_value"] for s_ in st] == [s_["f_value"] for s_ in runs[0][1]]
            for Fa, Fb in zip(F, runs[0][0]):
                for a, b in zip(Fa, Fb):
                    assert np.array_equal(a, b)
        F, st, lg = runs[-1]
        if arith == P.ARITH_F64:
            Fo, sto, logo, tol = oracle_problem(Xs[0], probs[0][1], init[0], lg[0], **kw)
            err, cgap = _check_problem(0, F[0], st[0], lg[0], Fo, sto, logo, tol)
            say("split problem vs oracle: factors %.2e (tolerance %.1e), cost %.2e" % (err, tol, cgap))
        else:
            Fo, sto, _ = o.cpd(Xs[0], probs[0][1], init[0], cost_form=o.COST_GRAM, **kw)
            assert st[0]["rel_cost"] == pytest.approx(sto["rel_cost"], abs=2e-6)


def test_batch_fp32_matrix_cores_vs_fp64(h):
    """arith = F32 runs the batch MTTKRPs on the FP32 matrix cores (FP32 products of the FP32
    tensor and the FP32-rounded Khatri-Rao values, FP64 totals every 12 k-steps); every
    other step stays FP64.  Each problem's final relative cost agrees with the FP64 run and
    with the oracle to 2e-6 absolute (measured on the 512-problem batch after 100 sweeps:
    at most 7.2e-7, median 9e-9); both block sizes give the same bits on different grids."""
    probs = _problems(12, seed=13)
    say("fp32 batch:", len(probs), "problems")
    Xs = _tensors(probs, seed0=70)
    init = [o.gen_factors(list(d), R, seed=300 + p) for p, (d, R) in enumerate(probs)]
    h.batch_set(Xs, [R for _, R in probs])
    kw = dict(max_iter=20, threshold_error=0.0)
    _, st64, _, _ = h.cpd_batch(init, block=512, grid=64, with_log=False, **kw)
    F32a, st32, _, _ = h.cpd_batch(init, block=512, grid=64, with_log=False, arith=P.ARITH_F32, **kw)
    F32b, st32b, _, _ = h.cpd_batch(init, block=512, grid=5, with_log=False, arith=P.ARITH_F32, **kw)
    gaps = []
    for p, ((dims, R), X) in enumerate(zip(probs, Xs)):
        Fo, sto, _ = o.cpd(X, R, init[p], cost_form=o.COST_GRAM, **kw)
        gaps.append(abs(st32[p]["rel_cost"] - st64[p]["rel_cost"]))
        assert st32[p]["rel_cost"] == pytest.approx(st64[p]["rel_cost"], abs=2e-6), p
        assert st32[p]["rel_cost"] == pytest.approx(sto["rel_cost"], abs=2e-6), p
        assert st32b[p]["f_value"] == st32[p]["f_value"], p
        for a, b in zip(F32a[p], F32b[p]):
            assert np.array_equal(a, b), p
    say("fp32 vs fp64 final relative cost gap: max %.3e" % max(gaps))
    with pytest.raises(P.PtnError):
        h.cpd_batch(init, arith=P.ARITH_BF16X3, with_log=False, **kw)


def test_batch_launch_shapes_agree_bitwise(h):
    """Each problem is solved by one block in a fixed order of operations, so the grid
    never changes a result: one block size on different grids (and the autotuned shape,
    for its block size) gives identical bits."""
    probs = _problems(20, seed=12)
    say("launch shapes:", len(probs), "problems")
    Xs = _tensors(probs, seed0=50)
    init = [o.gen_factors(list(d), R, seed=200 + p) for p, (d, R) in enumerate(probs)]
    h.batch_set(Xs, [R for _, R in probs])
    kw = dict(max_iter=8, threshold_error=0.0)
    runs = {}
    for shape in [(256, 1), (256, 7), (256, 512), (512, 3), (512, 256), (0, 0)]:
        F, st, _, rep = h.cpd_batch(init, block=shape[0], grid=shape[1], with_log=False, **kw)
        say("shape", shape, "->", (rep["block"], rep["grid"]), "seconds", rep["seconds"])
        runs[shape] = (rep["block"], F, [s["f_value"] for s in st])
    for bs in (256, 512):
        same = [v for v in runs.values() if v[0] == bs]
        for v in same[1:]:
            assert v[2] == same[0][2]
            for Fa, Fb in zip(v[1], same[0][1]):
                for a, b in zip(Fa, Fb):
                    assert np.array_equal(a, b)
    # the two block sizes sum the mode-0 partials and the cost terms over different
    # thread counts: equal to rounding
    a256 = [v for v in runs.values() if v[0] == 256][0]
    a512 = [v for v in runs.values() if v[0] == 512][0]
    assert a256[2] == pytest.approx(a512[2], rel=1e-11)


def test_batch_threshold_stop(h):
    """Problems stop on their own threshold test at the oracle's sweep and are left alone
    afterwards; the stopping sweep takes no line search."""
    probs = [((40, 30, 20), 4), ((60, 50, 40), 6), ((25, 25, 25), 3), ((70, 20, 30), 5)]
    Xs = _tensors(probs, snr_db=30.0, seed0=8)
    init = [o.gen_factors(list(d), R, seed=9 + p) for p, (d, R) in enumerate(probs)]
    h.batch_set(Xs, [R for _, R in probs])
    kw = dict(max_iter=120, threshold_error=0.05)
    Fg, stg, logg, _ = h.cpd_batch(init, block=256, grid=2, **kw)
    for p, ((dims, R), X) in enumerate(zip(probs, Xs)):
        Fo, sto, logo, tol = oracle_problem(X, R, init[p], logg[p], **kw)
        say("problem", p, "stops at", stg[p]["ao_iter"], "oracle", sto["ao_iter"])
        assert 1 < sto["ao_iter"] < 120
        assert logg[p][-1]["stopped"] == 1 and logg[p][-1]["accel_tried"] == 0
        _check_problem(p, Fg[p], stg[p], logg[p], Fo, sto, logo, tol)


def test_batch_no_accel_no_norm(h):
    probs = _problems(4, seed=13)[:8]
    Xs = _tensors(probs, seed0=70)
    init = [o.gen_factors(list(d), R, seed=300 + p) for p, (d, R) in enumerate(probs)]
    h.batch_set(Xs, [R for _, R in probs])
    kw = dict(max_iter=6, threshold_error=0.0, acceleration=0, normalization=0)
    Fg, stg, logg, _ = h.cpd_batch(init, block=512, grid=4, **kw)
    for p, ((dims, R), X) in enumerate(zip(probs, Xs)):
        Fo, sto, logo, tol = oracle_problem(X, R, init[p], logg[p], **kw)
        _check_problem(p, Fg[p], stg[p], logg[p], Fo, sto, logo, tol)


def test_batch_generator_and_init_match_oracle(h):
    probs = [((20, 30, 40), 5), ((128, 7, 64), 16), ((3, 128, 2), 1)]
    dims = [d for d, _ in probs]
    ranks = [R for _, R in probs]
    h.batch_generate(dims, ranks, seed=40, snr_db=25.0)
    for p, (d, R) in enumerate(probs):
        X, _ = o.make_tensor(list(d), R, seed=40 + p, snr_db=25.0)
        X32 = X.astype(np.float32).astype(np.float64)
        got = h.batch_tensor(p)
        assert P.rel_err(got, X32) < 1e-7
        assert np.mean(got == X32) > 0.999
    F = h.batch_init_factors(seed=7)
    for p, (d, R) in enumerate(probs):
        ref = o.gen_factors(list(d), R, seed=7 + p)
        for a, b in zip(F[p], ref):
            assert np.array_equal(a, b)


def test_batch_invalid_arguments(h):
    X = [np.ones((4, 5, 6), order="F")]
    with pytest.raises(P.PtnError):
        h.batch_set([np.ones((129, 2, 2), order="F")], [2])       # mode above 128
    with pytest.raises(P.PtnError):
        h.batch_set(X, [17])                                       # rank above 16
    h.batch_set(X, [3])
    init = [[np.ones((I, 3)) for I in (4, 5, 6)]]
    with pytest.raises(P.PtnError):
        h.cpd_batch(init, block=128, grid=4)                       # unsupported block size
    with pytest.raises(P.PtnError):
        h.cpd_batch(init, block=256, grid=0)                       # grid missing
    with pytest.raises(P.PtnError):
        h.cpd_batch(init, constraints=[0, 1, 1])                   # unconstrained mode
