"""Pins the oracle's MLP (FP64 forward/backward) with central finite differences and a numpy forward."""
import ctypes as C

import numpy as np
import pytest

import oracle

NIN, NH = 16, 64


def _params(rng):
    W1 = rng.normal(0, 0.4, (NH, NIN))
    W1[:, 12:] = 0.0
    return dict(W1=W1, b1=rng.normal(0, 0.1, NH), W2=rng.normal(0, 0.2, (NH, NH)), b2=rng.normal(0, 0.1, NH),
                W3=rng.normal(0, 0.3, NH), b3=rng.normal(0, 0.1, 1))


def _fwd(orc, X, p):
    Y = np.zeros(len(X))
    a = {k: np.ascontiguousarray(v, np.float64) for k, v in p.items()}
    orc.orc_mlp_forward_d(oracle.ptr(X), len(X), NIN, oracle.ptr(a["W1"]), oracle.ptr(a["b1"]), NH, oracle.ptr(a["W2"]),
                          oracle.ptr(a["b2"]), NH, oracle.ptr(a["W3"]), oracle.ptr(a["b3"]), oracle.ptr(Y))
    return Y


def _loss(orc, X, T, p):
    return 0.5 * np.mean((_fwd(orc, X, p) - T) ** 2)


def test_forward_matches_numpy(orc):
    rng = np.random.default_rng(0)
    p = _params(rng)
    X = np.ascontiguousarray(rng.normal(size=(50, NIN)))
    silu = lambda z: z / (1 + np.exp(-z))
    h1 = silu(X @ p["W1"].T + p["b1"])
    h2 = silu(h1 @ p["W2"].T + p["b2"])
    np.testing.assert_allclose(_fwd(orc, X, p), h2 @ p["W3"] + p["b3"][0], rtol=1e-12, atol=1e-12)


def test_backward_matches_central_differences(orc):
    rng = np.random.default_rng(1)
    p = _params(rng)
    X = np.ascontiguousarray(rng.normal(size=(40, NIN)))
    X[:, 12:] = 0.0
    T = np.ascontiguousarray(rng.normal(size=40))
    g = {k: np.zeros_like(v) for k, v in p.items()}
    loss = C.c_double()
    a = {k: np.ascontiguousarray(v, np.float64) for k, v in p.items()}
    orc.orc_mlp_backward_d(oracle.ptr(X), oracle.ptr(T), len(X), NIN, oracle.ptr(a["W1"]), oracle.ptr(a["b1"]), NH,
                           oracle.ptr(a["W2"]), oracle.ptr(a["b2"]), NH, oracle.ptr(a["W3"]), oracle.ptr(a["b3"]),
                           oracle.ptr(g["W1"]), oracle.ptr(g["b1"]), oracle.ptr(g["W2"]), oracle.ptr(g["b2"]),
                           oracle.ptr(g["W3"]), oracle.ptr(g["b3"]), C.byref(loss))
    assert loss.value == pytest.approx(_loss(orc, X, T, p), rel=1e-14)
    eps = 1e-6
    for name in p:
        flat = p[name].reshape(-1)
        idx = rng.choice(flat.size, min(25, flat.size), replace=False)
        for t in idx:
            if name == "W1" and (t % NIN) >= 12:
                continue
            old = flat[t]
            flat[t] = old + eps
            lp = _loss(orc, X, T, p)
            flat[t] = old - eps
            lm = _loss(orc, X, T, p)
            flat[t] = old
            fd = (lp - lm) / (2 * eps)
            an = g[name].reshape(-1)[t]
            assert abs(fd - an) <= 1e-4 * max(abs(fd), 1e-6) + 1e-9, (name, t, fd, an)
