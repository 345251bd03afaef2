"""Python handles of the fused solver objects of include/rx_sph_host.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import check, load


class PBFSolver:
    """rxPBDSPH_GPU (fused PBF step) for a Scene."""

    STATS = 64

    def __init__(self, scene, pos4: np.ndarray | None = None, vel4: np.ndarray | None = None):
        self.lib = load()
        self.scene = scene
        env = scene.env()
        self.h = self.lib.rxPBDSPH_GPU_Create(C.byref(env))
        if not self.h:
            check(self.lib)
            raise RuntimeError("rxPBDSPH_GPU_Create failed")
        if getattr(scene, "grad_mass", 0):
            self.option(_abi.RX_PBF_OPT_GRAD_MASS, 1.0)
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


class SPHSolver:
    """rxSPH_GPU (fused WCSPH step, hipGraph replay after the first step) for a Scene."""

    def __init__(self, scene, pos4: np.ndarray | None = None, vel4: np.ndarray | None = None, graph: bool = True):
        self.lib = load()
        self.scene = scene
        env = scene.env()
        self.h = self.lib.rxSPH_GPU_Create(C.byref(env))
        if not self.h:
            check(self.lib)
            raise RuntimeError("rxSPH_GPU_Create failed")
        self.lib.rxSPH_GPU_UseGraph(self.h, 1 if graph else 0)
        if pos4 is None:
            pos4, vel4 = scene.particles()
        pos4 = np.ascontiguousarray(pos4, np.float32)
        self.n = pos4.shape[0]
        v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        if self.lib.rxSPH_GPU_SetParticles(self.h, pos4.ctypes.data, None if v is None else v.ctypes.data, self.n):
            check(self.lib)

    def step(self, steps: int = 1, dt: float | None = None):
        if self.lib.rxSPH_GPU_Update(self.h, self.scene.dt if dt is None else dt, steps):
            check(self.lib)

    def sync(self):
        if self.lib.rxSPH_GPU_Sync(self.h):
            check(self.lib)

    def array(self, kind: int) -> np.ndarray:
        vec = kind in (_abi.RX_ARRAY_POSITION, _abi.RX_ARRAY_VELOCITY)
        out = np.zeros((self.n, 4) if vec else self.n, np.float32)
        if self.lib.rxSPH_GPU_GetArray(self.h, kind, out.ctypes.data, self.n):
            check(self.lib)
        return out

    def close(self):
        if self.h:
            self.lib.rxSPH_GPU_Destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IceSM:
    """rxIceSM: batched shape matching of one ice body (include/rx_ice_funcs.h)."""

    NBINS = 5

    def __init__(self, rest_pos4: np.ndarray, alpha=0.9, beta=0.3, dt=1.0 / 120.0, gravity=(0.0, -9.8, 0.0),
                 box_min=(-10.0, -10.0, -10.0), box_max=(10.0, 10.0, 10.0), iterations=2, volume_conservation=1,
                 restitution=0.9, cluster_radius=0.1, graph=True):
        self.lib = load()
        e = _abi.rxIceEnv()
        self.lib.rxIceEnvDefault(C.byref(e))
        e.alpha, e.beta, e.dt = alpha, beta, dt
        for k in range(3):
            e.gravity[k], e.box_min[k], e.box_max[k] = gravity[k], box_min[k], box_max[k]
        e.iterations, e.volume_conservation, e.restitution, e.cluster_radius = iterations, volume_conservation, restitution, cluster_radius
        self.env = e
        rest = np.ascontiguousarray(rest_pos4, np.float32)
        self.n = rest.shape[0]
        self.h = self.lib.rxIceSM_Create(C.byref(e), rest.ctypes.data, self.n)
        if not self.h:
            check(self.lib)
            raise RuntimeError("rxIceSM_Create failed")
        self.lib.rxIceSM_UseGraph(self.h, 1 if graph else 0)

    def set_state(self, pos4, vel4=None):
        p = np.ascontiguousarray(pos4, np.float32)
        v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
        if self.lib.rxIceSM_SetState(self.h, p.ctypes.data, None if v is None else v.ctypes.data):
            check(self.lib)

    def state(self):
        p = np.zeros((self.n, 4), np.float32)
        v = np.zeros((self.n, 4), np.float32)
        if self.lib.rxIceSM_GetState(self.h, p.ctypes.data, v.ctypes.data):
            check(self.lib)
        return p, v

    def step(self, steps: int = 1):
        if self.lib.rxIceSM_Update(self.h, steps):
            check(self.lib)

    def sync(self):
        if self.lib.rxIceSM_Sync(self.h):
            check(self.lib)

    def clusters(self):
        nc = C.c_int()
        nm = C.c_long()
        hist = np.zeros(128, np.int32)
        self.lib.rxIceSM_ClusterInfo(self.h, C.byref(nc), C.byref(nm), hist.ctypes.data, 128)
        start = np.zeros(nc.value + 1, np.int32)
        mem = np.zeros(max(nm.value, 1), np.int32)
        self.lib.rxIceSM_GetClusters(self.h, start.ctypes.data, mem.ctypes.data)
        return start, mem[: nm.value], hist

    def autotune(self):
        t = np.zeros(self.NBINS, np.int32)
        if self.lib.rxIceSM_Autotune(self.h, t.ctypes.data):
            check(self.lib)
        return t

    def set_tiles(self, tiles):
        t = np.ascontiguousarray(tiles, np.int32)
        if self.lib.rxIceSM_SetTiles(self.h, t.ctypes.data):
            check(self.lib)

    def set_active(self, active=None):
        """Melt mask (1 = ice) or None to remove it."""
        a = None if active is None else np.ascontiguousarray(active, np.int32)
        if self.lib.rxIceSM_SetActive(self.h, None if a is None else a.ctypes.data):
            check(self.lib)

    def kernel_times(self):
        a, b = C.c_float(), C.c_float()
        if self.lib.rxIceSM_KernelTimes(self.h, C.byref(a), C.byref(b)):
            check(self.lib)
        return a.value, b.value

    def close(self):
        if self.h:
            self.lib.rxIceSM_Destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Coupled:
    """rxCoupled: fluid + melting ice (include/rx_sph_host.h) for a CoupledScene."""

    def __init__(self, cs, data=None):
        self.lib = load()
        self.scene = cs
        pos, vel, temps, phase = cs.particles() if data is None else data
        env = cs.fluid.env()
        env.max_particles = cs.n
        ice = _abi.rxIceEnv()
        self.lib.rxIceEnvDefault(C.byref(ice))
        sp = cs.fluid.params()
        ice.alpha, ice.beta, ice.dt = cs.alpha, cs.beta, cs.fluid.dt
        for k in range(3):
            ice.gravity[k] = cs.fluid.gravity[k]
            ice.box_min[k] = sp.BoundaryMin[k] + sp.ParticleRadius
            ice.box_max[k] = sp.BoundaryMax[k] - sp.ParticleRadius
        ice.iterations, ice.volume_conservation, ice.restitution = cs.sm_iterations, 1, 0.9
        ice.cluster_radius = cs.radius_spacings * cs.fluid.spacing
        self.ice_env = ice
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        temps = np.ascontiguousarray(temps, np.float32)
        phase = np.ascontiguousarray(phase, np.int32)
        self.h = self.lib.rxCoupled_Create(C.byref(env), C.byref(ice), pos.ctypes.data, vel.ctypes.data, cs.n_fluid, cs.n_ice,
                                           temps.ctypes.data, phase.ctypes.data)
        if not self.h:
            check(self.lib)
            raise RuntimeError("rxCoupled_Create failed")
        self.n = cs.n
        self.option(_abi.RX_PBF_OPT_HEAT_TD, cs.heat_td)

    def option(self, opt: int, value: float):
        if self.lib.rxCoupled_SetOption(self.h, opt, value):
            check(self.lib)

    def set_network(self, net, warm_scale: float = 0.5, max_iters: int = 4, tol: float = 0.01):
        """The fluid's learned warm start (the network's weights are copied)."""
        if self.lib.rxCoupled_SetNetwork(self.h, net.h):
            check(self.lib)
        self.option(_abi.RX_PBF_OPT_WARM_SCALE, warm_scale)
        self.option(_abi.RX_PBF_OPT_MODE, _abi.RX_PBF_MODE_LEARNED)
        self.option(_abi.RX_PBF_OPT_ITERATIONS, max_iters)
        self.option(_abi.RX_PBF_OPT_ADAPTIVE, 1)
        self.option(_abi.RX_PBF_OPT_TOL, tol)

    def autotune(self):
        t = np.zeros(5, np.int32)
        if self.lib.rxCoupled_SetIceTiles(self.h, 1, t.ctypes.data):
            check(self.lib)
        return t

    def step(self, steps: int = 1):
        if self.lib.rxCoupled_Update(self.h, self.scene.fluid.dt, steps):
            check(self.lib)

    def sync(self):
        if self.lib.rxCoupled_Sync(self.h):
            check(self.lib)

    def array(self, kind: int) -> np.ndarray:
        vec = kind in (_abi.RX_ARRAY_POSITION, _abi.RX_ARRAY_VELOCITY)
        out = np.zeros((self.n, 4) if vec else self.n, np.float32)
        if self.lib.rxCoupled_GetArray(self.h, kind, out.ctypes.data, self.n):
            check(self.lib)
        return out

    def profile(self) -> dict:
        """PBF kernel classes: {class: (total ms, launches)} since the previous call (needs RX_PBF_OPT_PROFILE)."""
        ms = np.zeros(16, np.float32)
        cnt = np.zeros(16, np.int32)
        if self.lib.rxCoupled_GetProfile(self.h, ms.ctypes.data, cnt.ctypes.data, 16):
            check(self.lib)
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(_abi.PROF_CLASSES) if cnt[i]}

    def stats(self) -> np.ndarray:
        out = np.zeros(64, np.float32)
        if self.lib.rxCoupled_GetStats(self.h, out.ctypes.data, 64):
            check(self.lib)
        return out

    def close(self):
        if self.h:
            self.lib.rxCoupled_Destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
