"""CPU checks of the warm-start network's FP64 oracle: the back-propagated gradient
against central finite differences, and the feature layout on a known answer."""
import numpy as np

import mlp_oracle as m


def _init(rng, scale=1.0):
    w = np.zeros(m.N_PARAMS)
    o = 0
    for shape, fan in (((m.H, m.IN), m.IN), ((m.H,), 0), ((m.H, m.H), m.H), ((m.H,), 0), ((m.R_MAX, m.H), m.H),
                       ((m.R_MAX,), 0)):
        n = int(np.prod(shape))
        if fan:
            w[o:o + n] = rng.uniform(-1, 1, n) * np.sqrt(6.0 / fan) * scale
        else:
            w[o:o + n] = rng.uniform(-0.1, 0.1, n)
        o += n
    return w


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    n = 20
    X = rng.standard_normal((n, m.IN)) * 0.5
    T = rng.standard_normal((n, m.R_MAX))
    rank = rng.integers(1, m.R_MAX + 1, n)
    w = _init(rng)
    loss, g = m.loss_grad(w, X, T, rank)
    h = 1e-5
    offsets = [0, m.H * m.IN, m.H * m.IN + m.H, m.H * m.IN + m.H + m.H * m.H,
               m.H * m.IN + 2 * m.H + m.H * m.H, m.N_PARAMS - m.R_MAX]
    picks = []
    for a, b in zip(offsets, offsets[1:] + [m.N_PARAMS]):
        picks += list(rng.integers(a, b, 6))
    for i in picks:
        e = np.zeros_like(w)
        e[i] = h
        fd = (m.loss_grad(w + e, X, T, rank)[0] - m.loss_grad(w - e, X, T, rank)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(g[i]), 1e-3), (i, fd, g[i])
    # random directions through every parameter at once
    for _ in range(3):
        d = rng.standard_normal(m.N_PARAMS)
        d /= np.linalg.norm(d)
        fd = (m.loss_grad(w + h * d, X, T, rank)[0] - m.loss_grad(w - h * d, X, T, rank)[0]) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 * max(abs(g @ d), 1e-3)


def test_masked_entries_do_not_contribute():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((4, m.IN))
    T = rng.standard_normal((4, m.R_MAX))
    rank = np.array([3, 64, 1, 10])
    w = _init(rng)
    l0, g0 = m.loss_grad(w, X, T, rank)
    T2 = T.copy()
    T2[0, 3:] = 1e6                    # beyond rank 3: ignored
    l1, g1 = m.loss_grad(w, X, T2, rank)
    assert l0 == l1 and np.array_equal(g0, g1)


def test_feature_layout_known_answer():
    V = np.array([[1.0, 2.0], [2.0, 5.0]])
    M = np.array([[4.0, -8.0]])
    A = np.array([[0.5, 1.0]])
    X, T, scale = m.features(V, M, A, Anew=np.array([[2.0, 0.0]]))
    assert X.shape == (1, 2208)
    nz = {int(i): X[0, i] for i in np.nonzero(X[0])[0]}
    # vech: (0,0) -> 0, (0,1) -> 1, (1,1) -> 64; scaled by max diag 5
    assert nz == {0: 0.2, 1: 0.4, 64: 1.0, 2080: 0.5, 2081: -1.0, 2144: 0.5 * 5 / 8, 2145: 1.0 * 5 / 8}
    assert scale[0] == 8 / 5
    assert T[0, 0] == 2.0 * 5 / 8 and not T[0, 1:].any()
    assert m.vech_index(63, 63) == 2079


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38])
    r = m.bf16_round(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
