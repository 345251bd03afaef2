"""bf16-MFMA network kernels vs the FP64 oracle (weights and inputs rounded to bf16 first)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from fltk_sph_ice_amd import load
from fltk_sph_ice_amd.learned import Network, bf16_round, unpack_params

pytestmark = pytest.mark.gpu


def _oracle(orc, X, T, p):
    a = {k: np.ascontiguousarray(v, np.float64) for k, v in p.items()}
    Y = np.zeros(len(X))
    Xd = np.ascontiguousarray(X, np.float64)
    orc.orc_mlp_forward_d(oracle.ptr(Xd), len(X), 16, oracle.ptr(a["W1"]), oracle.ptr(a["b1"]), 64, oracle.ptr(a["W2"]),
                          oracle.ptr(a["b2"]), 64, oracle.ptr(a["W3"]), oracle.ptr(a["b3"]), oracle.ptr(Y))
    g = {k: np.zeros_like(v) for k, v in a.items()}
    loss = C.c_double()
    Td = np.ascontiguousarray(T, np.float64)
    orc.orc_mlp_backward_d(oracle.ptr(Xd), oracle.ptr(Td), len(X), 16, oracle.ptr(a["W1"]), oracle.ptr(a["b1"]), 64,
                           oracle.ptr(a["W2"]), oracle.ptr(a["b2"]), 64, oracle.ptr(a["W3"]), oracle.ptr(a["b3"]),
                           oracle.ptr(g["W1"]), oracle.ptr(g["b1"]), oracle.ptr(g["W2"]), oracle.ptr(g["b2"]),
                           oracle.ptr(g["W3"]), oracle.ptr(g["b3"]), C.byref(loss))
    return Y, g, loss.value


def _case(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 16)).astype(np.float32)
    X[:, 12:] = 0.0
    X = bf16_round(X)
    T = rng.normal(size=n).astype(np.float32)
    net = Network(seed=seed)
    p = unpack_params(net.params())
    # bf16 copies are what the MFMA kernels multiply; the oracle sees the same rounded weights
    p["W1"], p["W2"] = bf16_round(p["W1"]), bf16_round(p["W2"])
    return X, T, net, p


@pytest.mark.parametrize("n", [16, 100, 4096])
def test_forward_matches_oracle(orc, n):
    X, T, net, p = _case(n, 3)
    y = net.forward(X)
    yref, _, _ = _oracle(orc, X, T, p)
    # hidden activations are rounded to bf16 before the second product
    assert np.abs(y - yref).max() <= 2e-2 * max(np.abs(yref).max(), 1.0)
    assert np.corrcoef(y, yref)[0, 1] > 0.9999


@pytest.mark.parametrize("n", [64, 4096])
def test_gradients_match_oracle(orc, n):
    X, T, net, p = _case(n, 5)
    g, loss = net.gradients(X, T)
    _, gref, lref = _oracle(orc, X, T, p)
    assert loss == pytest.approx(lref, rel=2e-2)
    for k in gref:
        a, b = g[k].ravel(), gref[k].ravel()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        assert rel < 5e-2 and cos > 0.998, (k, rel, cos)


def test_training_reduces_loss():
    rng = np.random.default_rng(7)
    n = 20000
    X = rng.normal(size=(n, 16)).astype(np.float32)
    X[:, 12:] = 0.0
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3] ** 2).astype(np.float32)
    net = Network(seed=1)
    net.set_normalization(np.zeros(12), np.ones(12), 1.0, 0.0)
    hist = net.train(X, T, epochs=20, batch=1024, lr=3e-3)
    assert hist[-1] < 0.1 * hist[0], hist
    y = net.forward(X[:2000])
    assert np.mean((y - T[:2000]) ** 2) < 0.05


def test_training_is_bitwise_reproducible():
    # the gradient sums run over per-wavefront and per-workgroup partials in a fixed order (no float atomics),
    # so two trainings on the same data and seed end with the same parameters, bit for bit
    rng = np.random.default_rng(3)
    n = 8192
    X = rng.normal(size=(n, 16)).astype(np.float32)
    X[:, 12:] = 0.0
    T = (np.tanh(X[:, 0]) + 0.3 * X[:, 1] * X[:, 4]).astype(np.float32)
    params = []
    for _ in range(2):
        net = Network(seed=2)
        net.set_normalization(np.zeros(12), np.ones(12), 1.0, 0.0)
        net.train(X, T, epochs=3, batch=1024, lr=3e-3)
        params.append(net.params())
        net.close()
    assert np.array_equal(params[0].view(np.uint32), params[1].view(np.uint32))
