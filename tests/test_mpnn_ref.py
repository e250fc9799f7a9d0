"""The network's NumPy restatement (the checker of the HIP kernels) is pinned
by central finite differences of its float64 forward pass."""

import numpy as np

from oracle import mpnn_ref as M


def small_graph(n=40, seed=0):
    rng = np.random.default_rng(seed)
    # ring-plus-chords graph with symmetric neighbour lists of <= 3
    nb = -np.ones((n, 3), np.int64)
    for i in range(n):
        nb[i, 0] = (i + 1) % n
        nb[i, 1] = (i - 1) % n
    for i in range(0, n, 2):
        j = (i + n // 2) % n
        if nb[i, 2] < 0 and nb[j, 2] < 0:
            nb[i, 2], nb[j, 2] = j, i
    X = rng.normal(size=(n, M.F_IN))
    T = rng.normal(size=(n, M.N_T))
    return X, nb, T


def test_neighbour_mean_transpose_is_adjoint():
    X, nb, _ = small_graph()
    rng = np.random.default_rng(1)
    h, g = rng.normal(size=(40, 5)), rng.normal(size=(40, 5))
    assert np.isclose((M.neighbour_mean(h, nb) * g).sum(), (h * M.neighbour_mean_T(g, nb)).sum(), rtol=1e-12)


def test_gradients_match_finite_differences():
    X, nb, T = small_graph()
    p = {k: v.astype(np.float64) for k, v in M.init_params(3).items()}
    for k in p:
        if k.startswith("b"):
            p[k] = np.random.default_rng(len(k)).normal(0, 0.1, p[k].shape)
    loss, g, _ = M.loss_and_grad(p, X, nb, T)
    rng = np.random.default_rng(2)
    for k in p:
        idx = tuple(rng.integers(0, s) for s in p[k].shape)
        if k == "Wd" or k == "bd":
            idx = idx[:-1] + (int(rng.integers(0, M.N_T)),)
        e = 1e-6
        pp = {kk: v.copy() for kk, v in p.items()}
        pm = {kk: v.copy() for kk, v in p.items()}
        pp[k][idx] += e
        pm[k][idx] -= e
        fd = (M.loss_and_grad(pp, X, nb, T)[0] - M.loss_and_grad(pm, X, nb, T)[0]) / (2 * e)
        assert abs(fd - g[k][idx]) <= 1e-6 * max(1.0, abs(fd)), (k, idx, fd, g[k][idx])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7
    assert abs(r[3] + 3.140625) < 1e-7
