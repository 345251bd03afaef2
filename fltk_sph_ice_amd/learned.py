"""Learned warm start of the PBF solve: data generation with the solver itself, GPU training of the network,
and the learned solver.

The network predicts, from 12 features computed in the neighbour-list pass at the predicted positions x*,
the multiplier the solver converges to in this step (sum over the Jacobi iterations of lambda, scaled by
the rest-lattice denominator sd_ref).  The first position correction uses that prediction; the plain Jacobi
iterations follow until the max density error is at most the tolerance (the plain lambda pass is the judge).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import check, load
from .solvers import PBFSolver

NIN, NH = 16, 64
NFEAT = 12
FEATURE_NAMES = ["C", "lambda1", "sd_rel", "n_rel", "grad_mag", "grad_y", "C_start_nbr", "C_start_ring",
                 "C_start", "div_v", "wall", "v_y"]


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack_params(flat: np.ndarray) -> dict:
    o = 0
    out = {}
    for name, shape in (("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH,)), ("b3", (1,))):
        k = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + k], np.float64).reshape(shape)
        o += k
    return out


class Network:
    """rxMLP: 16 -> 64 -> 64 -> 1 SiLU network on bf16 MFMA."""

    def __init__(self, seed: int = 0):
        self.lib = load()
        self.h = self.lib.rxMLP_Create(seed)
        if not self.h:
            check(self.lib)
            raise RuntimeError("rxMLP_Create failed")
        self.np = self.lib.rxMLP_NumParams()

    def params(self) -> np.ndarray:
        p = np.zeros(self.np, np.float32)
        self.lib.rxMLP_GetParams(self.h, p.ctypes.data)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.np
        self.lib.rxMLP_SetParams(self.h, p.ctypes.data)

    def set_normalization(self, mu, sd, out_scale, out_bias):
        mu = np.ascontiguousarray(mu, np.float32)
        sd = np.ascontiguousarray(sd, np.float32)
        self.norm = (mu.copy(), sd.copy(), float(out_scale), float(out_bias))
        if self.lib.rxMLP_SetNormalization(self.h, mu.ctypes.data, sd.ctypes.data, out_scale, out_bias):
            check(self.lib)

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros(len(X), np.float32)
        if self.lib.rxMLP_Forward(self.h, X.ctypes.data, len(X), Y.ctypes.data):
            check(self.lib)
        return Y

    def gradients(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.np, np.float32)
        loss = C.c_float()
        if self.lib.rxMLP_Gradients(self.h, X.ctypes.data, T.ctypes.data, len(X), g.ctypes.data, C.byref(loss)):
            check(self.lib)
        return unpack_params(g), loss.value

    def train(self, X, T, epochs=30, batch=4096, lr=2e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        hist = np.zeros(epochs, np.float32)
        if self.lib.rxMLP_Train(self.h, X.ctypes.data, T.ctypes.data, len(X), epochs, batch, lr, seed, hist.ctypes.data):
            check(self.lib)
        return hist

    def close(self):
        if self.h:
            self.lib.rxMLP_Destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def feature_scales(solver: PBFSolver):
    sd, nref = C.c_float(), C.c_float()
    solver.lib.rxPBDSPH_GPU_FeatureScales(solver.h, C.byref(sd), C.byref(nref))
    return sd.value, nref.value


def generate_dataset(scenes, steps: int, samples_per_step: int, teacher_iters: int = 12, warmup: int = 5, seed: int = 0,
                     stride: int = 1):
    """(features, target) pairs from the solver itself: a teacher run of every scene (plain Jacobi with
    teacher_iters iterations) records, every `stride`-th of its `steps` steps, the features at x* and the
    multiplier the step applied in total (sum over the iterations of lambda, scaled by sd_ref)."""
    rng = np.random.default_rng(seed)
    Xs, Ts = [], []
    for sc in scenes:
        s = PBFSolver(sc)
        s.option(_abi.RX_PBF_OPT_MODE, _abi.RX_PBF_MODE_PLAIN)
        s.option(_abi.RX_PBF_OPT_ITERATIONS, teacher_iters)
        s.option(_abi.RX_PBF_OPT_RECORD_FEATURES, 1)
        sd_ref, _ = feature_scales(s)
        s.step(warmup)
        for t in range(steps):
            s.step(1)
            if (t + 1) % stride:
                continue
            F = np.zeros((s.n, NIN), np.float32)
            if s.lib.rxPBDSPH_GPU_GetArray(s.h, _abi.RX_ARRAY_FEATURES, F.ctypes.data, 16 * s.n):
                check(s.lib)
            lam = s.array(_abi.RX_ARRAY_LAMBDA_SUM)
            idx = rng.choice(s.n, min(samples_per_step, s.n), replace=False)
            Xs.append(F[idx])
            Ts.append(lam[idx] * sd_ref)
        s.close()
    return np.concatenate(Xs), np.concatenate(Ts)


def fit(X, T, epochs=40, batch=4096, lr=2e-3, seed=0) -> Network:
    mu = X[:, :NFEAT].mean(0)
    sd = X[:, :NFEAT].std(0) + 1e-6
    tmu, tsd = float(T.mean()), float(T.std() + 1e-12)
    net = Network(seed=seed)
    net.set_normalization(mu, sd, tsd, tmu)
    net.hist = net.train(X, (T - tmu) / tsd, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return net


def learned_solver(scene, net: Network, max_iters: int = 4, tol: float = 0.01, pos=None, vel=None,
                   warm_scale: float = 1.0) -> PBFSolver:
    s = PBFSolver(scene, pos, vel)
    s.option(_abi.RX_PBF_OPT_WARM_SCALE, warm_scale)
    if s.lib.rxPBDSPH_GPU_SetNetwork(s.h, net.h):
        check(s.lib)
    s.option(_abi.RX_PBF_OPT_MODE, _abi.RX_PBF_MODE_LEARNED)
    s.option(_abi.RX_PBF_OPT_ITERATIONS, max_iters)
    s.option(_abi.RX_PBF_OPT_ADAPTIVE, 1)
    s.option(_abi.RX_PBF_OPT_TOL, tol)
    return s
