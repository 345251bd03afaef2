"""Python handles of the fused solver objects of include/rx_sph_host.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import check, load


class PBFSolver:
    """rxPBDSPH_GPU (fused PBF step) for a Scene."""

    STATS = 32

    def __init__(self, scene, pos4: np.ndarray | None = None, vel4: np.ndarray | None = None):
        self.lib = load()
        self.scene = scene
        env = scene.env()
        self.h = self.lib.rxPBDSPH_GPU_Create(C.byref(env))
        if not self.h:
            check(self.lib)
            raise RuntimeError("rxPBDSPH_GPU_Create failed")
        if pos4 is None:
            pos4, vel4 = scene.particles()
        self.set_particles(pos4, vel4)

    def set_particles(self, pos4, vel4=None):
        pos4 = np.ascontiguousarray(pos4, np.float32)
        self.n = pos4.shape[0]
        v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        if self.lib.rxPBDSPH_GPU_SetParticles(self.h, pos4.ctypes.data, None if v is None else v.ctypes.data, self.n):
            check(self.lib)

    def option(self, opt: int, value: float):
        if self.lib.rxPBDSPH_GPU_SetOption(self.h, opt, value):
            check(self.lib)

    def step(self, steps: int = 1, dt: float | None = None):
        dt = self.scene.dt if dt is None else dt
        if self.lib.rxPBDSPH_GPU_Update(self.h, dt, steps):
            check(self.lib)

    def sync(self):
        if self.lib.rxPBDSPH_GPU_Sync(self.h):
            check(self.lib)

    def array(self, kind: int) -> np.ndarray:
        vec = kind in (_abi.RX_ARRAY_POSITION, _abi.RX_ARRAY_VELOCITY)
        out = np.zeros((self.n, 4) if vec else self.n, np.float32)
        if self.lib.rxPBDSPH_GPU_GetArray(self.h, kind, out.ctypes.data, self.n):
            check(self.lib)
        return out

    def positions(self):
        return self.array(_abi.RX_ARRAY_POSITION)

    def velocities(self):
        return self.array(_abi.RX_ARRAY_VELOCITY)

    def density(self):
        return self.array(_abi.RX_ARRAY_DENSITY)

    def stats(self) -> np.ndarray:
        out = np.zeros(self.STATS, np.float32)
        if self.lib.rxPBDSPH_GPU_GetStats(self.h, out.ctypes.data, self.STATS):
            check(self.lib)
        return out

    def profile(self) -> dict:
        """{kernel class: (total ms, launches)} since the previous call (needs RX_PBF_OPT_PROFILE)."""
        ms = np.zeros(16, np.float32)
        cnt = np.zeros(16, np.int32)
        if self.lib.rxPBDSPH_GPU_GetProfile(self.h, ms.ctypes.data, cnt.ctypes.data, 16):
            check(self.lib)
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(_abi.PROF_CLASSES) if cnt[i]}

    def close(self):
        if self.h:
            self.lib.rxPBDSPH_GPU_Destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
