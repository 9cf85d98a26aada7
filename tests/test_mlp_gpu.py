"""The NNLS warm-start network on the GPU: the BF16 forward and backward kernels against
the FP64 oracle (with the same BF16 operands), the harvested features against the
oracle's, the fallback invariant (a network whose proposals are rejected leaves the
decomposition bit-identical to the plain one), and what a trained network changes on a
held-out problem."""
import numpy as np
import pytest

import mlp_oracle as m
import oracle as o
import ptn_hip as P

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module")
def h():
    say("create handle")
    with P.Handle(0) as hd:
        yield hd


BLOCKS = [0, m.H * m.IN, m.H * m.IN + m.H, m.H * m.IN + m.H + m.H * m.H, m.H * m.IN + 2 * m.H + m.H * m.H,
          m.N_PARAMS - m.R_MAX, m.N_PARAMS]


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, m.IN)) * 0.3
    T = rng.standard_normal((n, m.R_MAX))
    rank = rng.integers(1, m.R_MAX + 1, n)
    return X, T, rank


def test_forward_matches_oracle(h):
    h.mlp_init(5)
    w = h.mlp_params().astype(np.float64)
    X, _, _ = _data(100, 0)
    Y = h.mlp_forward(X)
    Yo, _ = m.forward(w, X, bf16=True)
    say("forward rel err", P.rel_err(Y, Yo))
    # FP32 accumulation vs FP64 on the same BF16 operands; a hidden activation whose FP32 and
    # FP64 sums round to neighbouring BF16 values moves the output by ~1e-4 (measured 1.0e-4)
    assert P.rel_err(Y, Yo) < 5e-4


def test_gradient_matches_oracle(h):
    h.mlp_init(6)
    w = h.mlp_params().astype(np.float64)
    X, T, rank = _data(77, 1)                # not a multiple of 32: padded rows must not count
    loss, g = h.mlp_loss_grad(X, T, rank)
    lo, go = m.loss_grad(w, X, T, rank, bf16=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    for a, b in zip(BLOCKS, BLOCKS[1:]):
        err = P.rel_err(g[a:b].astype(np.float64), go[a:b])
        say("gradient block at", a, "rel err", err)
        assert err < 1e-6                    # measured <= 9e-8: FP32 accumulation only


def test_harvested_features_match_oracle(h):
    dims, R = [12, 10, 8], 3
    X, _ = o.make_tensor(dims, R, seed=3, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=4)
    h.set_tensor(X)
    h.mlp_data_clear()
    h.cpd(R, init, max_iter=1, threshold_error=0.0, mlp_harvest=1)
    assert h.mlp_data_size() == sum(dims)
    Xd, Td, rk = h.mlp_data_get(0, dims[0])      # the first update: mode 0 of sweep 1
    G = [o.gram(A) for A in init]
    V = G[1] * G[2]
    M = o.mttkrp(X, init, 0)
    Anew, _, _, _ = o.nesterov(V, M, init[0])
    Xo, To, _ = m.features(V, M, init[0], Anew)
    Xo = m.bf16_round(Xo)
    assert np.all(rk == R)
    assert np.mean(Xd == Xo) > 0.999 and np.max(np.abs(Xd - Xo)) <= 1e-2 * np.max(np.abs(Xo))
    assert P.rel_err(Td.astype(np.float64), To) < 1e-6


def test_rejected_proposals_leave_the_plain_result_bitwise(h):
    """A network whose outputs are NaN fails the finiteness check on every row: every NNLS
    starts from the previous row, exactly as without the network."""
    dims, R = [60, 50, 40], 6
    X, _ = o.make_tensor(dims, R, seed=8, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=9)
    h.set_tensor(X)
    w = np.zeros(m.N_PARAMS, dtype=np.float32)
    w[-m.R_MAX:] = np.nan
    h.mlp_set_params(w)
    for form in (P.MTTKRP_STANDARD, P.MTTKRP_DIMTREE):
        Fw, sw, lw = h.cpd(R, init, max_iter=12, threshold_error=0.0, warm_start=1, mttkrp_form=form)
        Fp, sp, lp = h.cpd(R, init, max_iter=12, threshold_error=0.0, mttkrp_form=form)
        assert all(sum(l["warm_rows"]) == 0 for l in lw)
        assert [l["nnls_iters"] for l in lw] == [l["nnls_iters"] for l in lp]
        assert sw["f_value"] == sp["f_value"]
        for a, b in zip(Fw, Fp):
            assert np.array_equal(a, b)
        # with refinement the non-finite proposals become the previous rows, which take the
        # refinement steps and are then judged; at a gate of 0 none passes (a zero residual
        # would need an exact solution), so the main solve is the plain one, and the counted
        # steps are the plain ones plus the refinement's
        Fr, sr, lr = h.cpd(R, init, max_iter=12, threshold_error=0.0, warm_start=1, warm_refine=5, warm_kkt_tol=0.0,
                           mttkrp_form=form)
        assert all(sum(l["warm_rows"]) == 0 for l in lr)
        assert all(0 <= s <= 5 for l in lr for s in l["warm_steps"]) and sum(map(sum, (l["warm_steps"] for l in lr)))
        assert [[a - b for a, b in zip(l["nnls_iters"], l["warm_steps"])] for l in lr] == \
            [l["nnls_iters"] for l in lp]
        assert sr["f_value"] == sp["f_value"]
        for a, b in zip(Fr, Fp):
            assert np.array_equal(a, b)


def test_trained_network_on_a_held_out_problem(h):
    """Harvest (inputs, Nesterov solution) pairs from two 100^3 R = 10 runs, train, and
    decompose a third tensor with and without the warm start.  Reported: NNLS steps in
    total, rows started from the network, and the sweep at which each run first reaches
    the plain run's final error.  Asserted: training lowers the loss; under the default
    gate (relative KKT residual <= 1e-6) a row takes the network's start only when the
    proposal is that close to the subproblem's solution, so the run ends at the plain
    solver's error (bit-identical when no row qualifies); under a loose gate the
    proposals are taken and the run stays finite."""
    dims, R = [100, 100, 100], 10
    h.mlp_data_clear()
    for seed in (0, 10):
        h.generate_tensor(dims, R, seed=seed, snr_db=30.0)
        h.cpd(R, h.generate_factors(dims, R, seed=seed + 1), max_iter=30, threshold_error=0.0, mlp_harvest=1)
    h.mlp_init(7)
    loss = h.mlp_train(20, batch=256, lr=1e-3, seed=1)
    say("training loss per epoch: first %.3e last %.3e, %d rows" % (loss[0], loss[-1], h.mlp_data_size()))
    assert loss[-1] < 0.2 * loss[0]
    h.generate_tensor(dims, R, seed=20, snr_db=30.0)
    init = h.generate_factors(dims, R, seed=21)
    Fp, sp, lp = h.cpd(R, init, max_iter=40, threshold_error=0.0)
    ip = sum(sum(l["nnls_iters"]) for l in lp)
    tgt = sp["rel_cost"] + 1e-9
    kp = next(i + 1 for i, l in enumerate(lp) if l["rel_cost"] <= tgt)
    for tol in (1e-6, 5e-2):
        Fw, sw, lw = h.cpd(R, init, max_iter=40, threshold_error=0.0, warm_start=1, warm_kkt_tol=tol)
        iw = sum(sum(l["nnls_iters"]) for l in lw)
        rows = sum(sum(l["warm_rows"]) for l in lw)
        kw = next((i + 1 for i, l in enumerate(lw) if l["rel_cost"] <= tgt), None)
        say("gate %g: NNLS steps plain %d, warm %d (saved %d); warm rows %d of %d; sweeps to the plain final "
            "error: plain %d, warm %s; final rel error plain %.12f warm %.12f"
            % (tol, ip, iw, ip - iw, rows, 40 * sum(dims), kp, kw, sp["rel_cost"], sw["rel_cost"]))
        if tol == 1e-6:
            assert abs(sw["rel_cost"] - sp["rel_cost"]) <= 1e-6 * sp["rel_cost"]
            if rows == 0:
                assert [l["nnls_iters"] for l in lw] == [l["nnls_iters"] for l in lp]
                for a, b in zip(Fw, Fp):
                    assert np.array_equal(a, b)
        else:
            assert rows > 0 and np.isfinite(sw["rel_cost"])


def test_warm_start_at_rank_64(h):
    """The warm start at its configuration's rank on a scaled tensor: pairs harvested from
    two 200^3 R = 64 runs (true rank 64, U[0,1) factors, SNR 15 dB), training, then a
    held-out 200^3 R = 64 problem decomposed with the plain start, the network's start
    under the 1e-6 KKT gate, the same after 10 refinement steps from every proposal, and
    the refined proposals under a loose 5e-2 gate.  Reported: rows that took the network's
    start, NNLS steps (refinement included) and the difference to the plain run, sweeps to
    the plain run's final error.  Asserted (the gate's promise): under the 1e-6 gate, with
    or without refinement, the final relative error is within 1e-6 of the plain run's."""
    dims, R, sweeps = [200, 200, 200], 64, 20
    h.mlp_data_clear()
    for seed in (30, 32):
        h.generate_tensor(dims, R, seed=seed, snr_db=15.0)
        h.cpd(R, h.generate_factors(dims, R, seed=seed + 1), max_iter=sweeps, threshold_error=0.0, mlp_harvest=1)
    pairs = h.mlp_data_size()
    h.mlp_init(11)
    loss = h.mlp_train(10, batch=512, lr=1e-3, seed=2)
    say("R=64: %d pairs, training loss %.3e -> %.3e" % (pairs, loss[0], loss[-1]))
    assert pairs == 2 * sweeps * sum(dims) and loss[-1] < loss[0]
    h.generate_tensor(dims, R, seed=34, snr_db=15.0)
    init = h.generate_factors(dims, R, seed=35)
    kw = dict(max_iter=sweeps, threshold_error=0.0)
    Fp, sp, lp = h.cpd(R, init, **kw)
    ip = sum(sum(l["nnls_iters"]) for l in lp)
    tgt = sp["rel_cost"] + 1e-9
    kp = next(i + 1 for i, l in enumerate(lp) if l["rel_cost"] <= tgt)
    for name, extra in (("gate 1e-6", dict(warm_kkt_tol=1e-6)),
                        ("refine 10, gate 1e-6", dict(warm_kkt_tol=1e-6, warm_refine=10)),
                        ("refine 10, gate 5e-2", dict(warm_kkt_tol=5e-2, warm_refine=10))):
        Fw, sw, lw = h.cpd(R, init, warm_start=1, **extra, **kw)
        iw = sum(sum(l["nnls_iters"]) for l in lw)
        rows = sum(sum(l["warm_rows"]) for l in lw)
        kw_ = next((i + 1 for i, l in enumerate(lw) if l["rel_cost"] <= tgt), None)
        say("R=64 %s: rows from the network %d of %d; NNLS steps plain %d, warm %d (saved %d); sweeps to the "
            "plain final error: plain %d, warm %s; final rel error plain %.12f warm %.12f"
            % (name, rows, sweeps * sum(dims), ip, iw, ip - iw, kp, kw_, sp["rel_cost"], sw["rel_cost"]))
        assert np.isfinite(sw["rel_cost"])
        if extra["warm_kkt_tol"] == 1e-6:
            assert abs(sw["rel_cost"] - sp["rel_cost"]) <= 1e-6


def test_warm_start_needs_a_network():
    with P.Handle(0) as h2:
        h2.set_tensor(np.ones((4, 5, 6), order="F"))
        with pytest.raises(P.PtnError):
            h2.cpd(2, [np.ones((I, 2)) for I in (4, 5, 6)], warm_start=1)
        h2.mlp_init(1)
        with pytest.raises(P.PtnError):
            h2.mlp_train(1, batch=32)             # empty training set
